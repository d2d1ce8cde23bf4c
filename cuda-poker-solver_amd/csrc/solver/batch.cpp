// Continuous batching over the engine's batch ABI (include/pps/batch.h).
//
// Threads: compile workers turn requests into compiled trees (host only)
// as the requests arrive; the calling thread owns the engine and runs the
// loop admit -> step -> check -> retire. Admission is in arrival order: a
// request that does not fit waits (and holds back later ones) until
// retirements free a slot or arena space.
#include "pps/batch.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>

#include "../host/compile.h"
#include "ppsg.h"

namespace pps {
namespace gpu {

namespace {

using Clock = std::chrono::steady_clock;

void check(int rc, const char* what) {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": " + ppsg_last_error());
}

// What a live spot keeps on the host once its tree is on the device.
struct Live {
  int req = -1;
  int slot = -1;
  int iterations = 0;
  double pot = 0.0, Z = 0.0;
  std::vector<double> w[2];
  int na0 = 0, nC0 = 0;  // root decide node (OOP)
  int64_t regOff0 = 0;
};

// A request after compilation.
struct Compiled {
  bool done = false;
  std::string error;
  std::unique_ptr<CompiledTree> tree;
  int64_t footprint = 0;
};

// Joins the compile workers on every exit path.
struct Workers {
  std::vector<std::thread> threads;
  std::atomic<bool> stop{false};
  std::condition_variable* cv = nullptr;
  ~Workers() {
    stop = true;
    if (cv) cv->notify_all();
    for (std::thread& t : threads)
      if (t.joinable()) t.join();
  }
};

struct Engine {
  ppsg_t* h = nullptr;
  ~Engine() {
    if (h) ppsg_destroy(h);
  }
};

}  // namespace

std::vector<SpotResult> serve(const std::vector<SpotRequest>& reqs, const ServeOptions& opt,
                              ServeReport* report) {
  const int n = (int)reqs.size();
  std::vector<SpotResult> out(n);
  ServeReport rep;
  if (n == 0) {
    if (report) *report = rep;
    return out;
  }
  if (opt.algo != "dcfr" && opt.algo != "dcfr-g2")
    throw std::runtime_error("serve: algo must be dcfr or dcfr-g2 (got '" + opt.algo + "')");
  if (opt.checkEvery <= 0 || opt.maxIters <= 0 || !(opt.targetFrac >= 0.0) || opt.maxSpots <= 0)
    throw std::runtime_error("serve: bad options");
  const int chunk = opt.checkEvery;
  const int cap = (opt.maxIters + chunk - 1) / chunk * chunk;

  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  if (opt.realTime)
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return reqs[a].arrivalMs < reqs[b].arrivalMs; });

  // one schedule table for every spot: dcfr and dcfr-g2 depend only on t
  std::vector<float> coef((size_t)cap * 3);
  for (int t = 0; t < cap; ++t) {
    double c[3];
    discountCoefs(t, cap, opt.algo, c);
    for (int k = 0; k < 3; ++k) coef[(size_t)t * 3 + k] = (float)c[k];
  }

  int64_t arena = opt.arenaBytes;
  if (arena <= 0) {
    int64_t fr = 0, tot = 0;
    check(ppsg_device_memory(&fr, &tot), "device memory query");
    arena = fr / 10 * 8;
  }
  Engine E;
  check(ppsg_batch_create(arena, opt.maxSpots, &E.h), "batch create");
  if (!opt.shapes.empty())
    check(ppsg_set_shapes(E.h, opt.shapes.data(), (int)opt.shapes.size()), "launch shapes");
  ppsg_t* h = E.h;
  // Terminal group size: a batch admitted up front shares each launch among
  // many spots, so larger groups (fewer, longer workgroups) pay off; a
  // stream keeps few spots live and wants the parallelism of small groups.
  int group = 4;
  if (opt.prefill) {
    const int together = std::min<int>(n, opt.maxLive > 0 ? opt.maxLive : opt.maxSpots);
    while (group < 32 && group * 8 < together) group *= 2;
  }
  check(ppsg_tune(h, PPSG_TUNE_GROUP, group), "tune");
  rep.groupSize = group;
  int addGroup = group;  // group size of the spots admitted next
  check(ppsg_tune(h, PPSG_TUNE_TERM_SPLIT, opt.termSplit), "terminal split");

  // ---- compile workers ----
  std::vector<Compiled> compiled(n);
  std::mutex mu;
  std::condition_variable cv;
  std::atomic<int> nextCompile{0}, admittedCount{0};
  const int lookahead = std::max(16, 2 * (opt.maxLive > 0 ? opt.maxLive : 32));
  const auto t0 = Clock::now();
  auto msNow = [&] { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); };
  Workers W;
  W.cv = &cv;
  auto work = [&] {
    for (;;) {
      const int k = nextCompile.fetch_add(1);
      if (k >= n) return;
      {
        // bound host memory: stay at most `lookahead` requests ahead of admission
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return W.stop || k - admittedCount.load() < lookahead; });
      }
      if (W.stop) return;
      const int r = order[k];
      if (opt.realTime) {
        const auto due = t0 + std::chrono::microseconds((int64_t)(reqs[r].arrivalMs * 1e3));
        while (!W.stop && Clock::now() < due)
          std::this_thread::sleep_for(std::min<Clock::duration>(due - Clock::now(),
                                                                std::chrono::milliseconds(20)));
        if (W.stop) return;
      }
      Compiled x;
      try {
        auto T = std::make_unique<CompiledTree>(compileSpot(reqs[r].spot, reqs[r].cfg, reqs[r].options.suitIsomorphism));
        if (T->empty) {
          x.error = "empty ranges";
        } else {
          const ppsg_tree_desc d = T->desc();
          if (ppsg_footprint(&d, &x.footprint) != 0) x.error = ppsg_last_error();
          else x.tree = std::move(T);
        }
      } catch (const std::exception& e) {
        x.error = e.what();
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        x.done = true;
        compiled[r] = std::move(x);
      }
      cv.notify_all();
    }
  };
  const int nThreads = std::max(1, std::min(opt.compileThreads, n));
  for (int i = 0; i < nThreads; ++i) W.threads.emplace_back(work);

  // ---- admit -> step -> check -> retire ----
  std::vector<Live> live;
  int admitted = 0, finished = 0;
  int64_t arenaCap = arena;
  auto isDone = [&](int r) {
    std::lock_guard<std::mutex> lk(mu);
    return compiled[r].done;
  };
  auto admitOne = [&](int r) -> bool {  // false: no room now
    Compiled& x = compiled[r];
    SpotResult& res = out[r];
    res.arrivalMs = opt.realTime ? reqs[r].arrivalMs : 0.0;
    if (!x.error.empty()) {
      res.error = x.error;
      ++finished;
      return true;
    }
    res.numNodes = x.tree->numNodes;
    res.footprint = x.footprint;
    if (x.footprint > arenaCap) {
      res.error = "spot needs " + std::to_string(x.footprint) + " device bytes; the arena holds " +
                  std::to_string(arenaCap);
      x.tree.reset();
      ++finished;
      return true;
    }
    if (!opt.prefill) {
      // a stream gives each spot the group size its live count calls for
      // when it is admitted (the same rule as a batch admitted at once);
      // the listed spots keep theirs, so nothing is re-cut
      int g = 4;
      while (g < 32 && g * 8 < (int)live.size() + 1) g *= 2;
      if (g != addGroup) {
        check(ppsg_tune(h, PPSG_TUNE_ADD_GROUP, g), "tune");
        addGroup = g;
      }
    }
    const ppsg_tree_desc d = x.tree->desc();
    int slot = -1;
    const double a0 = msNow();
    const int rc = ppsg_batch_add(h, &d, &slot);
    rep.admitMs += msNow() - a0;
    if (rc == -2) {
      if (!live.empty()) return false;
      res.error = std::string("cannot admit: ") + ppsg_last_error();
      x.tree.reset();
      ++finished;
      return true;
    }
    check(rc, "batch add");
    check(ppsg_batch_set_schedule(h, slot, coef.data(), cap), "set schedule");
    const CompiledTree& T = *x.tree;
    Live L;
    L.req = r;
    L.slot = slot;
    L.pot = (double)T.spot.pot;
    L.Z = pairMass(T);
    L.w[0] = T.w[0];
    L.w[1] = T.w[1];
    L.na0 = T.na[0];
    L.nC0 = T.listSize(0, 0);
    L.regOff0 = T.regOff[0];
    live.push_back(std::move(L));
    res.admitMs = msNow();
    x.tree.reset();
    return true;
  };

  std::vector<int> slots, which;
  std::vector<double> sums;
  std::vector<float*> noRows;  // walk_only: no root rows copied back
  // root value per pair of every wanted live spot: one walk of all live
  // spots, then the weighted root sums on the device (one launch, one copy
  // back; the same f64 values the host loop over copied rows gave)
  auto rootValues = [&](int tr, int mode, const std::vector<char>& want, std::vector<double>& v) {
    const int k = (int)live.size();
    slots.clear();
    which.clear();
    for (int i = 0; i < k; ++i)
      if (want[i] && live[i].Z > 0.0) {
        slots.push_back(live[i].slot);
        which.push_back(i);
      }
    // every live spot: one walk of all; fewer (the spots that retire): a
    // walk of just those (the others' items masked), the same rows for them
    if ((int)slots.size() == k)
      check(ppsg_batch_walk(h, tr, mode, 0, nullptr, nullptr), "batch walk");
    else if (!slots.empty()) {
      noRows.assign(slots.size(), nullptr);
      check(ppsg_batch_walk_only(h, tr, mode, (int)slots.size(), slots.data(), noRows.data()),
            "batch walk");
    }
    sums.assign(slots.size(), 0.0);
    check(ppsg_batch_root_sums(h, tr, (int)slots.size(), slots.data(), sums.data()), "root sums");
    v.assign(k, 0.0);
    for (size_t j = 0; j < which.size(); ++j) v[which[j]] = sums[j] / live[which[j]].Z;
  };

  while (finished < n) {
    // 1. admission, in arrival order (prefill: before the first step, wait
    // for every request until one does not fit)
    const bool filling = opt.prefill && rep.steps == 0;
    while (admitted < n && (opt.maxLive <= 0 || (int)live.size() < opt.maxLive)) {
      const int r = order[admitted];
      if (!isDone(r)) {
        if (!filling) break;
        std::unique_lock<std::mutex> lk(mu);
        cv.wait_for(lk, std::chrono::milliseconds(5), [&] { return compiled[r].done; });
        continue;
      }
      if (!admitOne(r)) break;
      ++admitted;
      admittedCount = admitted;
      cv.notify_all();
    }
    if (live.empty()) {
      if (finished >= n) break;
      std::unique_lock<std::mutex> lk(mu);
      const int r = admitted < n ? order[admitted] : -1;
      cv.wait_for(lk, std::chrono::milliseconds(5), [&] { return r < 0 || compiled[r].done; });
      continue;
    }
    rep.peakLive = std::max(rep.peakLive, (int)live.size());

    // 2. one step: `chunk` iterations of every live spot
    const double b0 = msNow();
    check(ppsg_batch_step(h, chunk), "batch step");
    double ms = 0.0;
    check(ppsg_last_iterate_ms(h, &ms), "step time");
    rep.deviceMs += ms;
    rep.stepLive.push_back((double)live.size());
    rep.stepDeviceMs.push_back(ms);
    rep.steps += 1;
    rep.iterations += (int64_t)chunk * (int64_t)live.size();
    for (Live& L : live) L.iterations += chunk;

    // 3. exploitability of every live spot (best-response walks)
    const double w0 = msNow();
    const std::vector<char> all(live.size(), 1);
    std::vector<double> br0, br1;
    rootValues(0, PPSG_MODE_BR, all, br0);
    rootValues(1, PPSG_MODE_BR, all, br1);
    std::vector<char> retire(live.size(), 0);
    std::vector<double> expl(live.size(), 0.0);
    bool any = false;
    for (size_t i = 0; i < live.size(); ++i) {
      expl[i] = (br0[i] + br1[i] - live[i].pot) / 2.0;
      retire[i] = expl[i] <= opt.targetFrac * live[i].pot || live[i].iterations >= cap;
      any |= retire[i] != 0;
    }
    // 4. retirement: final EVs and root strategy, then the slot is freed
    if (any) {
      std::vector<double> ev0, ev1;
      rootValues(0, PPSG_MODE_EV, retire, ev0);
      rootValues(1, PPSG_MODE_EV, retire, ev1);
      std::vector<Live> keep;
      for (size_t i = 0; i < live.size(); ++i) {
        Live& L = live[i];
        if (!retire[i]) {
          keep.push_back(std::move(L));
          continue;
        }
        SpotResult& res = out[L.req];
        res.ok = true;
        res.iterations = L.iterations;
        res.expl = expl[i];
        res.converged = expl[i] <= opt.targetFrac * L.pot;
        res.ev0 = ev0[i];
        res.ev1 = ev1[i];
        // the root's action rows sit PPSG_ROW_STRIDE(nC0) apart
        const int64_t ld = PPSG_ROW_STRIDE(L.nC0);
        std::vector<float> rows((size_t)((L.na0 - 1) * ld + L.nC0));
        check(ppsg_batch_read(h, L.slot, PPSG_STRAT, L.regOff0, (int64_t)rows.size(), rows.data()),
              "strategy read");
        for (int a = 1; a < L.na0; ++a)
          std::copy(rows.begin() + a * ld, rows.begin() + a * ld + L.nC0, rows.begin() + a * L.nC0);
        rows.resize((size_t)L.na0 * L.nC0);
        res.rootStrategy = rangeFrequencies(rows, L.w[0], L.na0, L.nC0);
        check(ppsg_batch_remove(h, L.slot), "batch remove");
        res.retireMs = msNow();
        ++finished;
      }
      live.swap(keep);
      int c = 0;
      const double c0 = msNow();
      // device-side compaction costs a few ms: drop retired spots at once
      check(ppsg_batch_compact(h, 0.0, &c), "compact");
      rep.compactMs += msNow() - c0;
      rep.compactions += c;
    }
    rep.walkMs += msNow() - w0;
    rep.stepCheckMs.push_back(msNow() - w0);
    rep.busyMs += msNow() - b0;
  }
  rep.wallMs = msNow();
  if (report) *report = rep;
  return out;
}

BatchProfile profileBatch(const std::vector<SpotRequest>& reqs, const std::string& algo, int iters,
                          int64_t arenaBytes, bool detail, int warmup) {
  if (algo != "dcfr" && algo != "dcfr-g2")
    throw std::runtime_error("profile: algo must be dcfr or dcfr-g2");
  if (iters <= 0 || reqs.empty()) throw std::runtime_error("profile: nothing to run");
  if (warmup < 0) throw std::runtime_error("profile: negative warmup");
  // warmup + replay + launch timing (+ attribution)
  const int tab = warmup + (detail ? 3 : 2) * iters;
  std::vector<float> coef((size_t)tab * 3);
  for (int t = 0; t < tab; ++t) {
    double c[3];
    discountCoefs(t, tab, algo, c);
    for (int k = 0; k < 3; ++k) coef[(size_t)t * 3 + k] = (float)c[k];
  }
  if (arenaBytes <= 0) {
    int64_t fr = 0, tot = 0;
    check(ppsg_device_memory(&fr, &tot), "device memory query");
    arenaBytes = fr / 10 * 8;
  }
  Engine E;
  check(ppsg_batch_create(arenaBytes, (int)reqs.size(), &E.h), "batch create");
  BatchProfile P;
  P.arenaBytes = arenaBytes;
  std::vector<CompiledTree> trees(reqs.size());
  std::vector<std::thread> th;
  const int nt = std::max(1, std::min<int>(8, (int)reqs.size()));
  std::atomic<int> next{0};
  std::vector<std::string> errs(reqs.size());
  for (int i = 0; i < nt; ++i)
    th.emplace_back([&] {
      for (int k; (k = next.fetch_add(1)) < (int)reqs.size();) {
        try {
          trees[k] = compileSpot(reqs[k].spot, reqs[k].cfg, reqs[k].options.suitIsomorphism);
        } catch (const std::exception& e) {
          errs[k] = e.what();
        }
      }
    });
  for (std::thread& t : th) t.join();
  for (size_t k = 0; k < reqs.size(); ++k) {
    if (!errs[k].empty()) throw std::runtime_error("profile: spot " + std::to_string(k) + ": " + errs[k]);
    if (trees[k].empty) continue;
    const ppsg_tree_desc d = trees[k].desc();
    int slot = -1;
    check(ppsg_batch_add(E.h, &d, &slot), "profile: batch add");
    check(ppsg_batch_set_schedule(E.h, slot, coef.data(), tab), "profile: schedule");
    trees[k] = CompiledTree();
    ++P.spots;
  }
  P.tuneMs.assign(kNumGroupCandidates, 0.0);
  check(ppsg_autotune(E.h, kGroupCandidates, kNumGroupCandidates, 2, &P.groupSize, P.tuneMs.data()),
        "profile: autotune");
  P.shapes.assign(1 + 6 * 1024, 0);
  check(ppsg_autotune_shapes(E.h, 3, P.shapes.data(), (int)P.shapes.size()), "profile: shapes");
  P.shapes.resize(1 + 6 * (size_t)P.shapes[0]);
  P.splitMs.assign(kNumSplitCandidates, 0.0);
  check(ppsg_autotune_param(E.h, PPSG_TUNE_TERM_SPLIT, kSplitCandidates, kNumSplitCandidates, 2,
                            &P.termSplit, P.splitMs.data()),
        "profile: terminal split");
  if (warmup > 0) check(ppsg_batch_step(E.h, warmup), "profile: warmup");
  check(ppsg_batch_step(E.h, iters), "profile: replay");
  double ms = 0.0;
  check(ppsg_last_iterate_ms(E.h, &ms), "profile: replay time");
  P.graphMsPerIter = ms / iters;
  double o[17] = {0};
  check(ppsg_profile(E.h, iters, o, 13), "profile: launches");
  P.fwdMs = o[0];
  P.fwdLaunches = o[1];
  P.bwdMs = o[2];
  P.bwdLaunches = o[3];
  P.iters = iters;
  P.foldMs = o[9] / iters;
  P.showdownMs = o[11] / iters;
  P.foldLaunches = o[10];
  P.showdownLaunches = o[12];
  if (detail) {
    check(ppsg_profile(E.h, iters, o, 17), "profile: attribution");
    P.decideMs = o[5] / iters;
    P.chanceMs = o[7] / iters;
    P.fwdDecideMs = o[13] / iters;
    P.fwdChanceMs = o[15] / iters;
  }
  int64_t info[6] = {0};
  check(ppsg_info(E.h, info, 6), "profile: info");
  P.kernelsPerIter = info[0];
  P.itemsPerIter = info[2];
  P.fwdBytes = (double)info[4];
  P.bwdBytes = (double)info[5];
  check(ppsg_kind_bytes(E.h, P.kindBytes), "profile: bytes by kind");
  return P;
}

}  // namespace gpu
}  // namespace pps

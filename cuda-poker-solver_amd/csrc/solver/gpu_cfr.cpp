// GpuPostflopSolver (include/pps/gpu_cfr.h) on top of the host compiler
// and the MI355X engine's C ABI. Mirrors the reference API's semantics
// (cuda/gpu_cfr.cu:1829-2853); host reductions are f64 over the f32 rows.
#include "pps/gpu_cfr.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

#include "../host/compile.h"
#include "solution_file.h"
#include "ppsg.h"

namespace pps {
namespace gpu {

namespace {

void check(int rc) {
  if (rc != 0) throw std::runtime_error(std::string("ppsg: ") + ppsg_last_error());
}

bool knownAlgo(const std::string& a) { return a == "dcfr" || a == "hs30" || a == "dcfr-g2"; }

std::vector<float> scheduleTable(int64_t from, int64_t to, int planned, const std::string& algo) {
  std::vector<float> tab((size_t)to * 3);
  double c0[3];
  discountCoefs((int)from, planned, algo, c0);
  for (int64_t t = 0; t < to; ++t) {
    double c[3];
    if (t < from) {
      std::copy(c0, c0 + 3, c);  // never read again: the counter only grows
    } else {
      discountCoefs((int)t, planned, algo, c);
    }
    for (int k = 0; k < 3; ++k) tab[(size_t)t * 3 + k] = (float)c[k];
  }
  return tab;
}

}  // namespace

struct GpuPostflopSolver::Impl {
  CompiledTree T;
  ppsg_t* eng = nullptr;
  int slot = 0;           // the spot's slot in eng (a one-spot engine: 0)
  bool attached = false;  // eng is a SpotServer's arena, shared with other spots
  std::string lastAlgo;
  double Z = 0.0;
  double lastMs = -1.0;
  // rows read from a solution file before the spot is attached to an arena
  std::vector<float> pendingReg, pendingStr;
  int64_t pendingIters = -1;
  std::vector<int32_t> subEnd;  // per node: one past its last descendant (pre-order ids)

  // EV / BR walk of node u's subtree only (its ancestors' forward launches
  // give it its reach); u's subtree rows equal a full walk's.
  void walkSubtree(int tr, int mode, int32_t u) {
    const int64_t N = T.numNodes;
    if (subEnd.empty()) {
      subEnd.resize((size_t)N);
      for (int64_t v = N - 1; v >= 0; --v) {
        int32_t e = (int32_t)v + 1;
        const int64_t b = T.childBase[v], end = v + 1 < N ? T.childBase[v + 1] : (int64_t)T.childFlat.size();
        for (int64_t k = b; k < end; ++k) e = std::max(e, subEnd[(size_t)T.childFlat[k]]);
        subEnd[(size_t)v] = e;
      }
    }
    std::vector<int64_t> path;
    for (int32_t a = T.parent[u]; a >= 0; a = T.parent[a]) path.push_back(a);
    check(ppsg_batch_walk_range(eng, slot, tr, mode, u, subEnd[(size_t)u], path.data(),
                                (int)path.size()));
  }

  ~Impl() {
    if (!eng) return;
    if (attached) (void)ppsg_batch_remove(eng, slot);
    else ppsg_destroy(eng);
  }

  std::vector<float> read(int buf, int64_t off, int64_t n) const {
    std::vector<float> v((size_t)n);
    check(ppsg_batch_read(eng, slot, buf, off, n, v.data()));
    return v;
  }
  void write(int buf, int64_t off, int64_t n, const float* v) const {
    check(ppsg_batch_write(eng, slot, buf, off, n, v));
  }
  // Whole regret / strategy tables in node order: each decide node's
  // action-major block (na rows of nC floats, back to back) after the
  // blocks of all lower node ids (the row layout of the reference's tables
  // and of solution files). On the device the blocks sit breadth-first
  // (T.regOff) with each action row padded to T.regStride(u).
  std::vector<float> readTable(int buf) const {
    const std::vector<float> dev = read(buf, 0, T.regDevTotal);
    std::vector<float> out((size_t)T.regTotal);
    int64_t at = 0;
    for (const int32_t u : T.decideNodes) {
      const int64_t nC = T.listSize(T.owner[u], u), ld = T.regStride(u);
      for (int a = 0; a < T.na[u]; ++a, at += nC)
        std::copy(dev.begin() + T.regOff[u] + a * ld, dev.begin() + T.regOff[u] + a * ld + nC,
                  out.begin() + at);
    }
    return out;
  }
  void writeTable(int buf, const float* v) const {
    std::vector<float> dev((size_t)T.regDevTotal, 0.0f);
    int64_t at = 0;
    for (const int32_t u : T.decideNodes) {
      const int64_t nC = T.listSize(T.owner[u], u), ld = T.regStride(u);
      for (int a = 0; a < T.na[u]; ++a, at += nC)
        std::copy(v + at, v + at + nC, dev.begin() + T.regOff[u] + a * ld);
    }
    write(buf, 0, T.regDevTotal, dev.data());
  }
  // One decide node's block, action-major with nC floats per action.
  std::vector<float> readBlock(int buf, int64_t u) const {
    const int na = T.na[u];
    const int64_t nC = T.listSize(T.owner[u], (int)u), ld = T.regStride(u);
    std::vector<float> dev = read(buf, T.regOff[u], (na - 1) * ld + nC);
    for (int a = 1; a < na; ++a)
      std::copy(dev.begin() + a * ld, dev.begin() + a * ld + nC, dev.begin() + a * nC);
    dev.resize((size_t)(na * nC));
    return dev;
  }
  // EV / BR walk of this spot alone; cfv (n[tr] floats) may be null.
  void walk(int tr, int mode, float* cfv) const {
    if (!attached) {
      check(ppsg_walk(eng, tr, mode, cfv));
      return;
    }
    float* outs[1] = {cfv};
    check(ppsg_batch_walk_only(eng, tr, mode, 1, &slot, outs));
  }
  void setSchedule(const std::vector<float>& tab, int64_t n) const {
    check(ppsg_batch_set_schedule(eng, slot, tab.data(), n));
  }
  void setCounter(int64_t t) const { check(ppsg_batch_set_counter(eng, slot, t)); }
  int64_t counter() const {
    int64_t t = 0;
    check(ppsg_batch_get_counter(eng, slot, &t));
    return t;
  }
  // Calls that iterate or time a whole engine belong to one-spot engines.
  void ownEngine(const char* what) const {
    if (attached)
      throw std::runtime_error(std::string(what) +
                               ": this spot is served by a SpotServer; run it through the server");
  }

  // Range-weighted root value of traverser tr (EV or BR walk) / pair mass.
  double rootValue(int tr, int mode) {
    std::vector<float> cfv(T.n[tr]);
    walk(tr, mode, cfv.data());
    double v = 0.0;
    for (int i = 0; i < T.n[tr]; ++i) v += T.w[tr][i] * (double)cfv[i];
    return Z > 0.0 ? v / Z : 0.0;
  }

  double explNow() {
    const double v0 = rootValue(0, PPSG_MODE_BR);
    const double v1 = rootValue(1, PPSG_MODE_BR);
    return (v0 + v1 - (double)T.spot.pot) / 2.0;
  }

  const std::vector<int32_t>& listBase() const { return T.listBase[0]; }

  // node-local slot of base slot `base` in player p's list at node u, or -1
  int localSlot(int p, int u, int base) const {
    const int cl = T.cls[u];
    const int64_t off = T.clsListOff[p][cl];
    const int n = T.clsN[p][cl];
    const int32_t* b = T.listBase[p].data() + off;
    const int32_t* it = std::lower_bound(b, b + n, base);
    return (it != b + n && *it == base) ? (int)(it - b) : -1;
  }

  // Opponent mass valid against each of tr's combos at u, from layout tr's
  // reach row (per-card sums + identical-combo correction, f64).
  std::vector<double> massAt(int tr, int u, std::vector<float>* reachOut = nullptr) const {
    const int op = 1 - tr, cl = T.cls[u];
    const int nO = T.clsN[op][cl], nT = T.clsN[tr][cl];
    std::vector<float> reach = read(tr == 0 ? PPSG_REACH0 : PPSG_REACH1, T.reachOff[tr][u], nO);
    const int64_t offO = T.clsListOff[op][cl], offT = T.clsListOff[tr][cl];
    double perCard[52] = {0.0}, tot = 0.0;
    for (int j = 0; j < nO; ++j) {
      const int b = T.listBase[op][offO + j];
      tot += reach[j];
      perCard[T.cards[op][2 * b]] += reach[j];
      perCard[T.cards[op][2 * b + 1]] += reach[j];
    }
    std::vector<double> m(nT);
    for (int i = 0; i < nT; ++i) {
      const int b = T.listBase[tr][offT + i];
      double v = tot - perCard[T.cards[tr][2 * b]] - perCard[T.cards[tr][2 * b + 1]];
      const int so = T.sameOther[tr][b];
      if (so >= 0) {
        const int s = localSlot(op, u, so);
        if (s >= 0) v += reach[s];
      }
      m[i] = v;
    }
    if (reachOut) *reachOut = std::move(reach);
    return m;
  }

  std::vector<ActionLabel> labels(int u) const {
    std::vector<ActionLabel> out;
    for (int a = 0; a < T.na[u]; ++a) {
      const pf::TreeAction& act = T.actFlat[T.actBase[u] + a];
      out.push_back({act.kind, act.amount});
    }
    return out;
  }

  int decideIndexCheck(int idx) const {
    if (idx < 0 || idx >= (int)T.decideNodes.size())
      throw std::runtime_error("decide node index " + std::to_string(idx) + " out of range (" +
                               std::to_string(T.decideNodes.size()) + " decide nodes)");
    return T.decideNodes[idx];
  }

  int64_t childCount(int u) const {
    const int64_t end = u + 1 < T.numNodes ? T.childBase[u + 1] : (int64_t)T.childFlat.size();
    return end - T.childBase[u];
  }
};

GpuPostflopSolver::GpuPostflopSolver(const PostflopSpot& spot, const pf::BetConfig& cfg)
    : impl_(new Impl()) {
  try {
    init(spot, cfg);
  } catch (...) {
    delete impl_;
    impl_ = nullptr;
    throw;
  }
}

GpuPostflopSolver::GpuPostflopSolver(const PostflopSpot& spot, const pf::BetConfig& cfg,
                                     const SolverOptions& opt)
    : impl_(new Impl()) {
  try {
    init(spot, cfg, true, opt);
  } catch (...) {
    delete impl_;
    impl_ = nullptr;
    throw;
  }
}

GpuPostflopSolver::GpuPostflopSolver(Deferred, const PostflopSpot& spot, const pf::BetConfig& cfg,
                                     const SolverOptions& opt)
    : impl_(new Impl()) {
  try {
    init(spot, cfg, false, opt);
  } catch (...) {
    delete impl_;
    impl_ = nullptr;
    throw;
  }
}

GpuPostflopSolver::~GpuPostflopSolver() { delete impl_; }

// ---- server hooks ----

bool GpuPostflopSolver::beginSolve(int iterations, const std::string& algo) {
  Impl* I = impl_;
  iterationsRun_ = 0;
  if (I->T.empty) return false;
  if (!knownAlgo(algo))
    throw std::runtime_error("unknown algo '" + algo + "' (expected dcfr, dcfr-g2 or hs30)");
  reset();
  if (iterations <= 0) return false;
  I->lastAlgo = algo;
  I->lastMs = 0.0;
  I->setSchedule(scheduleTable(0, iterations, iterations, algo), iterations);
  return true;
}

bool GpuPostflopSolver::beginContinue(int iterations) {
  Impl* I = impl_;
  iterationsRun_ = 0;
  if (I->T.empty) return false;
  if (totalIters_ == 0) throw std::runtime_error("continueSolve: no prior solve() to continue");
  if (I->lastAlgo == "hs30" || !knownAlgo(I->lastAlgo))
    throw std::runtime_error("continueSolve: the last solve used '" + I->lastAlgo +
                             "', whose schedule depends on the planned iteration count; "
                             "continuation is defined for DCFR only");
  if (iterations <= 0) return false;
  const int64_t t0 = totalIters_;
  if (t0 + iterations > (int64_t)std::numeric_limits<int>::max())
    throw std::runtime_error("continueSolve: schedule index overflow");
  const int64_t tEnd = t0 + iterations;
  I->lastMs = 0.0;
  I->setSchedule(scheduleTable(t0, tEnd, (int)tEnd, I->lastAlgo), tEnd);
  return true;
}

void GpuPostflopSolver::advanced(int n, double deviceMs) {
  iterationsRun_ += n;
  totalIters_ += n;
  impl_->lastMs += deviceMs;
}

double GpuPostflopSolver::explFromSums(double s0, double s1) const {
  const Impl* I = impl_;
  const double v0 = I->Z > 0.0 ? s0 / I->Z : 0.0, v1 = I->Z > 0.0 ? s1 / I->Z : 0.0;
  return (v0 + v1 - (double)I->T.spot.pot) / 2.0;
}

int GpuPostflopSolver::slot() const { return impl_->T.empty ? -1 : impl_->slot; }

ppsg_t* GpuPostflopSolver::engineHandle() const { return impl_->T.empty ? nullptr : impl_->eng; }


void GpuPostflopSolver::init(const PostflopSpot& spot, const pf::BetConfig& cfg, bool device,
                             const SolverOptions& opt) {
  Impl* I = impl_;
  I->T = compileSpot(spot, cfg, opt.suitIsomorphism);
  if (I->T.empty) return;
  numNodes_ = (int)I->T.numNodes;
  maxDepth_ = I->T.maxDepth;
  I->Z = pairMass(I->T);
  if (!device) return;
  const ppsg_tree_desc d = I->T.desc();
  check(ppsg_create(&d, &I->eng));
}

void GpuPostflopSolver::attach(ppsg_t* arena) {
  Impl* I = impl_;
  if (I->T.empty || I->eng) return;
  const ppsg_tree_desc d = I->T.desc();
  int slot = -1;
  const int rc = ppsg_batch_add(arena, &d, &slot);
  if (rc == -2) throw ArenaFull("no room for this spot in the server's device arena");
  check(rc);
  I->eng = arena;
  I->slot = slot;
  I->attached = true;
  check(ppsg_batch_set_active(arena, slot, 0));  // parked until a job runs it
  if (I->pendingIters >= 0) {  // a loaded solution's rows
    I->writeTable(PPSG_REGRET, I->pendingReg.data());
    I->writeTable(PPSG_STRAT, I->pendingStr.data());
    I->setCounter(I->pendingIters);
    I->pendingReg = std::vector<float>();
    I->pendingStr = std::vector<float>();
    I->pendingIters = -1;
  }
}

void GpuPostflopSolver::reset() {
  iterationsRun_ = 0;
  totalIters_ = 0;
  if (impl_->T.empty) return;
  check(ppsg_batch_reset(impl_->eng, impl_->slot));
}

// GPU_CFR_PROFILE set: up to 64 of the iterations run launch by launch with
// per-depth timings on stderr instead of the graph replay (the reference's
// profiling mode, cuda/gpu_cfr.cu:1895-1987); only those iterations count.
void GpuPostflopSolver::profileDepthsRun(int iterations) {
  Impl* I = impl_;
  const int pit = std::min(iterations, 64);
  const int D = I->T.maxDepth + 1;
  std::vector<double> out((size_t)(1 + 4 * D + 4), 0.0);
  check(ppsg_profile_depths(I->eng, pit, out.data(), (int)out.size()));
  const int Dk = (int)out[0];
  std::vector<int> kinds((size_t)D * 4, 0);  // per depth: decide, showdown, fold, chance
  for (int64_t u = 0; u < I->T.numNodes; ++u) {
    const int k = I->T.kind[u];
    const int col = k == PPSG_DECIDE ? 0 : k == PPSG_SHOWDOWN ? 1 : k == PPSG_FOLD ? 2 : 3;
    kinds[(size_t)I->T.depth[u] * 4 + col]++;
  }
  double fwd = 0.0, bwd = 0.0;
  for (int d = 0; d < std::min(D, Dk); ++d) {
    const double* t = out.data() + 1 + 4 * d;
    fwd += t[0] + t[2];
    bwd += t[1] + t[3];
    std::fprintf(stderr,
                 "gpu_cfr: d%-2d decide=%-5d show=%-5d fold=%-5d chance=%-3d  fwd %6.1f/%6.1f us"
                 "  bwd %6.1f/%6.1f us\n",
                 d, kinds[(size_t)d * 4], kinds[(size_t)d * 4 + 1], kinds[(size_t)d * 4 + 2],
                 kinds[(size_t)d * 4 + 3], t[0] / pit, t[2] / pit, t[1] / pit, t[3] / pit);
  }
  const double* term = out.data() + 1 + 4 * Dk;
  std::fprintf(stderr, "gpu_cfr: terminals fold %6.1f/%6.1f us  showdown %6.1f/%6.1f us\n",
               term[0] / pit, term[1] / pit, term[2] / pit, term[3] / pit);
  bwd += term[0] + term[1] + term[2] + term[3];
  std::fprintf(stderr, "gpu_cfr: per-iter (serialized) fwd %.1f us  bwd %.1f us over %d iters\n",
               fwd / pit, bwd / pit, pit);
  iterationsRun_ = pit;
  totalIters_ += pit;
  I->lastMs = (fwd + bwd) / 1e3;
}

void GpuPostflopSolver::replayChunks(int iterations, double target) {
  Impl* I = impl_;
  I->ownEngine("solve");
  if (std::getenv("GPU_CFR_PROFILE")) {
    profileDepthsRun(iterations);
    return;
  }
  const int chunk = kCheckEvery;
  int done = 0;
  double ms = 0.0;
  while (done < iterations) {
    const int n = std::min(chunk, iterations - done);
    check(ppsg_iterate(I->eng, n));
    double m = 0.0;
    check(ppsg_last_iterate_ms(I->eng, &m));
    ms += m;
    done += n;
    iterationsRun_ = done;
    totalIters_ += n;
    if (target > 0.0 && done < iterations && I->explNow() <= target) break;
  }
  I->lastMs = ms;
}

void GpuPostflopSolver::solve(int iterations, const std::string& algo, double target) {
  Impl* I = impl_;
  iterationsRun_ = 0;
  if (I->T.empty) return;
  if (!knownAlgo(algo))
    throw std::runtime_error("unknown algo '" + algo + "' (expected dcfr, dcfr-g2 or hs30)");
  reset();
  if (iterations <= 0) return;
  I->lastAlgo = algo;
  I->setSchedule(scheduleTable(0, iterations, iterations, algo), iterations);
  replayChunks(iterations, target);
}

void GpuPostflopSolver::continueSolve(int iterations, double target) {
  Impl* I = impl_;
  iterationsRun_ = 0;
  if (I->T.empty) return;
  if (totalIters_ == 0) throw std::runtime_error("continueSolve: no prior solve() to continue");
  if (I->lastAlgo == "hs30" || !knownAlgo(I->lastAlgo))
    throw std::runtime_error("continueSolve: the last solve used '" + I->lastAlgo +
                             "', whose schedule depends on the planned iteration count; "
                             "continuation is defined for DCFR only");
  if (iterations <= 0) return;
  const int64_t t0 = totalIters_;
  if (t0 + iterations > (int64_t)std::numeric_limits<int>::max())
    throw std::runtime_error("continueSolve: schedule index overflow");
  const int64_t tEnd = t0 + iterations;
  I->setSchedule(scheduleTable(t0, tEnd, (int)tEnd, I->lastAlgo), tEnd);
  replayChunks(iterations, target);
}

pf::NodeStats GpuPostflopSolver::stats() {
  Impl* I = impl_;
  pf::NodeStats st{};
  if (I->T.empty || !(I->Z > 0.0)) return st;
  if (const char* sm = std::getenv("GPU_CFR_SEED_STRAT")) {
    // probe profile of cuda/gpu_cfr.cu:2062-2097
    const int mode = std::atoi(sm);
    std::vector<float> rows((size_t)std::max<int64_t>(I->T.regDevTotal, 0), 0.0f);
    for (size_t mi = 0; mi < I->T.decideNodes.size(); ++mi) {
      const int u = I->T.decideNodes[mi];
      const int dpt = I->T.depth[u];
      bool pick;
      switch (mode) {
        case 1: pick = true; break;
        case 2: pick = dpt <= 2; break;
        case 3: pick = dpt > 2; break;
        case 4: pick = dpt == 0; break;
        case 5: pick = dpt == 1; break;
        case 6: pick = dpt == 2; break;
        case 7: pick = mi == 1; break;
        case 8: pick = mi == 2; break;
        case 9: pick = dpt == 1 && mi != 1; break;
        default: pick = false;
      }
      if (!pick || I->T.na[u] < 2) continue;
      const int nC = I->T.listSize(I->T.owner[u], u);
      for (int c = 0; c < nC; ++c) {
        const float p0 = (float)(0.15 + 0.7 * ((c % 8) / 7.0));
        rows[I->T.regOff[u] + c] = p0;
        rows[I->T.regOff[u] + I->T.regStride(u) + c] = 1.0f - p0;
      }
    }
    if (!rows.empty()) I->write(PPSG_STRAT, 0, (int64_t)rows.size(), rows.data());
    std::fprintf(stderr, "seeded strat rows (mode %d)\n", mode);
  }
  st.ev0 = I->rootValue(0, PPSG_MODE_EV);
  st.ev1 = I->rootValue(1, PPSG_MODE_EV);
  const double v0 = I->rootValue(0, PPSG_MODE_BR);
  const double v1 = I->rootValue(1, PPSG_MODE_BR);
  st.expl = (v0 + v1 - (double)I->T.spot.pot) / 2.0;
  st.pairMass = (int64_t)(I->Z * (1LL << 20));
  return st;
}

std::vector<double> GpuPostflopSolver::nodeStrategy(int decideIdx) {
  Impl* I = impl_;
  if (I->T.empty || decideIdx < 0 || decideIdx >= (int)I->T.decideNodes.size()) return {};
  const int u = I->T.decideNodes[decideIdx];
  if (I->T.clsNBoard[I->T.cls[u]] != I->T.spot.nBoard) return {};
  const int p = I->T.owner[u], na = I->T.na[u], nC = I->T.listSize(p, u);
  return rangeFrequencies(I->readBlock(PPSG_STRAT, u), I->T.w[p], na, nC);
}

std::vector<double> GpuPostflopSolver::rootStrategy() {
  Impl* I = impl_;
  if (I->T.empty) return {};
  const int na = I->T.na[0], nC = I->T.listSize(0, 0);
  return rangeFrequencies(I->readBlock(PPSG_STRAT, 0), I->T.w[0], na, nC);
}

int GpuPostflopSolver::numDecideNodes() const { return (int)impl_->T.decideNodes.size(); }

DecideNodeInfo GpuPostflopSolver::decideNode(int decideIdx) const {
  const Impl* I = impl_;
  const int u = I->decideIndexCheck(decideIdx);
  DecideNodeInfo out;
  out.nodeId = u;
  out.player = I->T.owner[u];
  out.depth = I->T.depth[u];
  out.nBoard = I->T.clsNBoard[I->T.cls[u]];
  out.nCombos = I->T.listSize(out.player, u);
  out.actions = I->labels(u);
  for (int a = 0; a < I->T.na[u]; ++a) out.children.push_back(I->T.childFlat[I->T.childBase[u] + a]);
  return out;
}

ComboStrategy GpuPostflopSolver::comboStrategy(int decideIdx) {
  Impl* I = impl_;
  if (I->T.empty) throw std::runtime_error("solver compiled to empty ranges");
  const int u = I->decideIndexCheck(decideIdx);
  const int p = I->T.owner[u], na = I->T.na[u], nC = I->T.listSize(p, u);
  ComboStrategy out;
  out.actions = I->labels(u);
  const int64_t off = I->T.clsListOff[p][I->T.cls[u]];
  out.combos.assign(I->T.listBase[p].begin() + off, I->T.listBase[p].begin() + off + nC);
  const std::vector<float> rows = I->readBlock(PPSG_STRAT, u);
  out.freqs.assign((size_t)na * nC, 0.0);
  for (int c = 0; c < nC; ++c) {
    double s = 0.0;
    for (int a = 0; a < na; ++a) s += rows[(size_t)a * nC + c];
    for (int a = 0; a < na; ++a)
      out.freqs[(size_t)a * nC + c] = s > 0.0 ? (double)rows[(size_t)a * nC + c] / s : 1.0 / na;
  }
  return out;
}

ComboEv GpuPostflopSolver::rootEvPerCombo() {
  Impl* I = impl_;
  ComboEv out;
  if (I->T.empty) return out;
  for (int tr = 0; tr < 2; ++tr) {
    std::vector<float> cfv(I->T.n[tr]);
    I->walk(tr, PPSG_MODE_EV, cfv.data());
    const int op = 1 - tr;
    double perCard[52] = {0.0}, tot = 0.0;
    for (int j = 0; j < I->T.n[op]; ++j) {
      tot += I->T.w[op][j];
      perCard[I->T.cards[op][2 * j]] += I->T.w[op][j];
      perCard[I->T.cards[op][2 * j + 1]] += I->T.w[op][j];
    }
    for (int i = 0; i < I->T.n[tr]; ++i) {
      double m = tot - perCard[I->T.cards[tr][2 * i]] - perCard[I->T.cards[tr][2 * i + 1]];
      if (I->T.sameOther[tr][i] >= 0) m += I->T.w[op][I->T.sameOther[tr][i]];
      out.combos[tr].push_back(i);
      out.mass[tr].push_back(m);
      out.ev[tr].push_back(m > 0.0 ? (double)cfv[i] / m : 0.0);
    }
  }
  return out;
}

NodeEv GpuPostflopSolver::nodeEv(int decideIdx) {
  Impl* I = impl_;
  if (I->T.empty) throw std::runtime_error("solver compiled to empty ranges");
  const int u = I->decideIndexCheck(decideIdx);
  const int na = I->T.na[u], dec = I->T.owner[u];
  NodeEv out;
  out.decider = dec;
  out.actions = I->labels(u);
  out.actionEv.assign(na, {});
  for (int tr = 0; tr < 2; ++tr) {
    I->walkSubtree(tr, PPSG_MODE_EV, u);
    NodeEv::Side& S = out.side[tr];
    const int cl = I->T.cls[u];
    const int nC = I->T.clsN[tr][cl];
    const int64_t off = I->T.clsListOff[tr][cl];
    S.combos.assign(I->T.listBase[tr].begin() + off, I->T.listBase[tr].begin() + off + nC);
    S.mass = I->massAt(tr, u);
    const std::vector<float> row = I->read(PPSG_CFV, I->T.cfvOff[u], nC);
    S.ev.assign(nC, 0.0);
    for (int i = 0; i < nC; ++i) S.ev[i] = S.mass[i] > 0.0 ? (double)row[i] / S.mass[i] : 0.0;
    const std::vector<double>& wSelf = I->T.w[tr];
    if (tr == dec) {
      const std::vector<float> srows = I->readBlock(PPSG_STRAT, u);
      for (int a = 0; a < na; ++a) {
        const int ch = I->T.childFlat[I->T.childBase[u] + a];
        const int nChild = I->T.listSize(tr, ch);
        const std::vector<float> crow = I->read(PPSG_CFV, I->T.cfvOff[ch], nChild);
        std::vector<double>& pc = out.actionEv[a].perCombo;
        pc.assign(nC, 0.0);
        for (int i = 0; i < nC; ++i) {
          if (S.mass[i] <= 0.0) continue;
          const int slot = I->T.kind[ch] == PPSG_CHANCE ? I->localSlot(tr, ch, S.combos[i]) : i;
          if (slot >= 0) pc[i] = (double)crow[slot] / S.mass[i];
        }
        double num = 0.0, den = 0.0, num0 = 0.0, den0 = 0.0;
        for (int i = 0; i < nC; ++i) {
          if (S.mass[i] <= 0.0) continue;
          double s = 0.0;
          for (int b = 0; b < na; ++b) s += srows[(size_t)b * nC + i];
          const double wb = wSelf[S.combos[i]] * S.mass[i];
          num0 += wb * pc[i];
          den0 += wb;
          if (s <= 0.0) continue;
          const double wt = wb * (double)srows[(size_t)a * nC + i] / s;
          num += wt * pc[i];
          den += wt;
        }
        out.actionEv[a].aggEv = den > 0.0 ? num / den : (den0 > 0.0 ? num0 / den0 : 0.0);
      }
      for (int i = 0; i < nC; ++i) {
        double s = 0.0;
        for (int a = 0; a < na; ++a) s += srows[(size_t)a * nC + i];
        if (s <= 0.0) continue;
        double ev = 0.0;
        for (int a = 0; a < na; ++a)
          ev += (double)srows[(size_t)a * nC + i] / s * out.actionEv[a].perCombo[i];
        S.ev[i] = ev;
      }
    }
    double num = 0.0, den = 0.0;
    for (int i = 0; i < nC; ++i) {
      if (S.mass[i] <= 0.0) continue;
      const double wt = wSelf[S.combos[i]] * S.mass[i];
      num += wt * S.ev[i];
      den += wt;
    }
    S.aggEv = den > 0.0 ? num / den : 0.0;
  }
  return out;
}

RunoutSummary GpuPostflopSolver::runoutSummary(int decideIdx, int actionIdx) {
  Impl* I = impl_;
  if (I->T.empty) throw std::runtime_error("solver compiled to empty ranges");
  const int u = I->decideIndexCheck(decideIdx);
  if (actionIdx < 0 || actionIdx >= I->T.na[u])
    throw std::runtime_error("action index " + std::to_string(actionIdx) + " out of range");
  const int ch = I->T.childFlat[I->T.childBase[u] + actionIdx];
  if (I->T.kind[ch] != PPSG_CHANCE) throw std::runtime_error("action does not lead to a deal");
  RunoutSummary out;
  // every dealt card, in card order; under suit isomorphism a card mapped to
  // another is reported with that card's subtree (its aggregates are the
  // same by symmetry)
  const int dcl = I->T.cls[ch];
  const int nb = I->T.dealN[dcl];
  const int64_t doff = I->T.dealOff[dcl];
  // first decide node under a branch (breadth-first through further deals)
  auto firstDecide = [&](int start) -> int {
    if (I->T.kind[start] == PPSG_DECIDE) return start;
    std::vector<int> q{start};
    for (size_t qi = 0; qi < q.size(); ++qi) {
      const int x = q[qi];
      for (int64_t k = 0; k < I->childCount(x); ++k) {
        const int c = I->T.childFlat[I->T.childBase[x] + k];
        if (I->T.kind[c] == PPSG_DECIDE) return c;
        if (I->T.kind[c] == PPSG_CHANCE) q.push_back(c);
      }
    }
    return -1;
  };
  out.branches.resize(nb);
  int firstIdx = -1;
  for (int b = 0; b < nb; ++b) {
    const int bc = I->T.childFlat[I->T.childBase[ch] + I->T.dealSib[doff + b]];
    out.branches[b].card = I->T.dealCard[doff + b];
    const int fd = firstDecide(bc);
    if (fd >= 0) {
      out.branches[b].nextDecide = I->T.decideOf[fd];
      if (firstIdx < 0) firstIdx = I->T.decideOf[fd];
    }
  }
  if (firstIdx < 0) return out;
  const int fu = I->T.decideNodes[firstIdx];
  out.decider = I->T.owner[fu];
  out.actions = I->labels(fu);
  for (int tr = 0; tr < 2; ++tr) {
    I->walkSubtree(tr, PPSG_MODE_EV, ch);
    const std::vector<double>& wSelf = I->T.w[tr];
    for (int b = 0; b < nb; ++b) {
      const int bi = out.branches[b].nextDecide;
      if (bi < 0) continue;
      const int f = I->T.decideNodes[bi];
      const int cl = I->T.cls[f];
      const int nC = I->T.clsN[tr][cl], nOpp = I->T.clsN[1 - tr][cl];
      const int64_t offC = I->T.clsListOff[tr][cl];
      std::vector<float> reach;
      const std::vector<double> mass = I->massAt(tr, f, &reach);
      const std::vector<float> row = I->read(PPSG_CFV, I->T.cfvOff[f], nC);
      double num = 0.0, den = 0.0;
      for (int i = 0; i < nC; ++i) {
        if (mass[i] <= 0.0) continue;
        const double wt = wSelf[I->T.listBase[tr][offC + i]] * mass[i];
        num += wt * (double)row[i] / mass[i];
        den += wt;
      }
      out.branches[b].aggEv[tr] = den > 0.0 ? num / den : 0.0;
      const int bna = I->T.na[f];
      if (I->T.owner[f] != tr && bna == (int)out.actions.size()) {
        const std::vector<float> srows = I->readBlock(PPSG_STRAT, f);
        std::vector<double>& freq = out.branches[b].freq;
        freq.assign(bna, 0.0);
        double z = 0.0;
        for (int c = 0; c < nOpp; ++c) {
          double s = 0.0;
          for (int a = 0; a < bna; ++a) s += srows[(size_t)a * nOpp + c];
          if (s <= 0.0) continue;
          z += reach[c];
          for (int a = 0; a < bna; ++a) freq[a] += reach[c] * (double)srows[(size_t)a * nOpp + c] / s;
        }
        if (z > 0.0)
          for (double& f2 : freq) f2 /= z;
      }
    }
  }
  return out;
}

std::vector<uint8_t> GpuPostflopSolver::playerCards(int player) const {
  if (player < 0 || player > 1) throw std::runtime_error("player must be 0 (OOP) or 1 (IP)");
  return impl_->T.cards[player];
}

int GpuPostflopSolver::numBaseCombos(int player) const {
  if (player < 0 || player > 1) throw std::runtime_error("player must be 0 (OOP) or 1 (IP)");
  return impl_->T.n[player];
}

std::vector<double> GpuPostflopSolver::playerWeights(int player) const {
  if (player < 0 || player > 1) throw std::runtime_error("player must be 0 (OOP) or 1 (IP)");
  return impl_->T.w[player];
}

TreeStructure GpuPostflopSolver::treeStructure() const {
  const CompiledTree& T = impl_->T;
  TreeStructure t;
  t.kinds = T.kind;
  t.childBase = T.childBase;
  t.children = T.childFlat;
  t.boards.reserve((size_t)T.numNodes * 5);
  t.nBoard.reserve((size_t)T.numNodes);
  for (int64_t u = 0; u < T.numNodes; ++u) {
    const int cl = T.cls[u];
    for (int i = 0; i < 5; ++i) t.boards.push_back(i < T.clsNBoard[cl] ? T.clsBoard[cl][i] : 0);
    t.nBoard.push_back((uint8_t)T.clsNBoard[cl]);
  }
  return t;
}

double GpuPostflopSolver::iterationBytes() const { return impl_->T.iterationBytes(); }

std::vector<float> GpuPostflopSolver::rows(int which) const {
  if (impl_->T.empty) return {};
  return impl_->readTable(which == 0 ? PPSG_REGRET : PPSG_STRAT);
}

void GpuPostflopSolver::setRows(int which, const std::vector<float>& v) {
  if (impl_->T.empty) return;
  if ((int64_t)v.size() != impl_->T.regTotal) throw std::runtime_error("row size mismatch");
  impl_->writeTable(which == 0 ? PPSG_REGRET : PPSG_STRAT, v.data());
}

void GpuPostflopSolver::halfStep(int tr, int t, const std::string& algo, int planned) {
  Impl* I = impl_;
  if (I->T.empty) return;
  if (!knownAlgo(algo)) throw std::runtime_error("unknown algo '" + algo + "'");
  if (tr != 0 && tr != 1) throw std::runtime_error("traverser must be 0 or 1");
  I->ownEngine("halfStep");
  const int64_t saved = I->counter();
  I->setSchedule(scheduleTable(t, (int64_t)t + 1, planned, algo), (int64_t)t + 1);
  I->setCounter(t);
  check(ppsg_half_step(I->eng, tr));
  I->setCounter(saved);
}

std::vector<double> GpuPostflopSolver::profileIterations(int iters, bool detail) {
  Impl* I = impl_;
  if (I->T.empty || iters <= 0) return {};
  if (totalIters_ == 0 || I->lastAlgo == "hs30")
    throw std::runtime_error("profileIterations: needs a prior dcfr / dcfr-g2 solve");
  I->ownEngine("profileIterations");
  const int64_t t0 = totalIters_, tEnd = t0 + iters;
  I->setSchedule(scheduleTable(t0, tEnd, (int)tEnd, I->lastAlgo), tEnd);
  std::vector<double> out(detail ? 17 : 13, 0.0);
  check(ppsg_profile(I->eng, iters, out.data(), (int)out.size()));
  // detail mode puts the rows and the schedule counter back: no iteration
  // is added
  if (!detail) totalIters_ = tEnd;
  return out;
}

std::vector<int64_t> GpuPostflopSolver::engineInfo() const {
  std::vector<int64_t> out(6, 0);
  if (impl_->T.empty) return out;
  impl_->ownEngine("engineInfo");
  check(ppsg_info(impl_->eng, out.data(), 6));
  return out;
}

double GpuPostflopSolver::rootValue(int tr, int mode, std::vector<float>* rootCfv) {
  Impl* I = impl_;
  if (I->T.empty) return 0.0;
  if (tr != 0 && tr != 1) throw std::runtime_error("traverser must be 0 or 1");
  std::vector<float> cfv(I->T.n[tr]);
  I->walk(tr, mode, cfv.data());
  double v = 0.0;
  for (int i = 0; i < I->T.n[tr]; ++i) v += I->T.w[tr][i] * (double)cfv[i];
  if (rootCfv) *rootCfv = cfv;
  return I->Z > 0.0 ? v / I->Z : 0.0;
}

double GpuPostflopSolver::lastSolveDeviceMs() const { return impl_->lastMs; }

// ---------------------------------------------------------------------------
// Solution files (solver/solution_file.h): the spot, the bet abstraction and
// the rows in node order. Loading recompiles the tree from the stored spot
// and configuration; the row count has to match it.
// ---------------------------------------------------------------------------

void GpuPostflopSolver::save(const std::string& path) const {
  const Impl* I = impl_;
  if (I->T.empty) throw std::runtime_error("save: the spot compiled to an empty tree (no rows)");
  if (I->T.suitIso)
    throw std::runtime_error(
        "save: a suit-isomorphic solve has its own tree, and a solution file records no "
        "isomorphism option (it would reload as the full tree)");
  Solution s;
  s.spot = I->T.spot;
  s.cfg = I->T.cfg;
  s.nodes = I->T.numNodes;
  s.iterations = totalIters_;
  s.algo = I->lastAlgo;
  s.rows = I->T.regTotal;
  s.regret = I->readTable(PPSG_REGRET);
  s.strat = I->readTable(PPSG_STRAT);
  writeSolution(path, s);
}

GpuPostflopSolver::GpuPostflopSolver(const std::string& solutionPath) : impl_(new Impl()) {
  try {
    loadFrom(solutionPath, true);
  } catch (...) {
    delete impl_;
    impl_ = nullptr;
    throw;
  }
}

GpuPostflopSolver::GpuPostflopSolver(Deferred, const std::string& solutionPath) : impl_(new Impl()) {
  try {
    loadFrom(solutionPath, false);
  } catch (...) {
    delete impl_;
    impl_ = nullptr;
    throw;
  }
}

void GpuPostflopSolver::loadFrom(const std::string& solutionPath, bool device) {
  Solution s = readSolution(solutionPath);
  init(s.spot, s.cfg, device);
  Impl* I = impl_;
  if (I->T.empty)
    throw std::runtime_error("load '" + solutionPath +
                             "': its spot compiles to an empty tree (a range is empty or "
                             "blocked by the board)");
  if (s.rows != I->T.regTotal)
    throw std::runtime_error("load '" + solutionPath + "': " + std::to_string(s.rows) +
                             " row entries, but the tree compiled from its spot has " +
                             std::to_string(I->T.regTotal) +
                             " (written by a tree builder with other rules or bet abstraction)");
  if (device) {
    I->writeTable(PPSG_REGRET, s.regret.data());
    I->writeTable(PPSG_STRAT, s.strat.data());
    I->setCounter(s.iterations);
  } else {
    I->pendingReg = std::move(s.regret);
    I->pendingStr = std::move(s.strat);
    I->pendingIters = s.iterations;
  }
  I->lastAlgo = s.algo;
  totalIters_ = s.iterations;
}

void GpuPostflopSolver::debugDump() {
  Impl* I = impl_;
  if (I->T.empty) return;
  for (int which = 0; which < 2; ++which) {
    // node order: the device table's row padding holds values no one reads
    const std::vector<float> v = I->readTable(which == 0 ? PPSG_STRAT : PPSG_REGRET);
    long long bad = 0;
    for (float x : v) bad += !std::isfinite(x);
    std::fprintf(stderr, "%s nonfinite=%lld of %lld\n", which == 0 ? "strat" : "regret", bad,
                 (long long)I->T.regTotal);
  }
}

}  // namespace gpu
}  // namespace pps

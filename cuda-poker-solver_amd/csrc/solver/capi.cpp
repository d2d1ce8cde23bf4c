// C ABI over GpuPostflopSolver and the integer layer (include/pps_capi.h).
// Exceptions never cross the boundary: every entry returns a status and
// leaves the message in a thread-local buffer.
#include "pps_capi.h"

#include <algorithm>
#include <cstring>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "host/compile.h"
#include "pps/batch.h"
#include "pps/cards.h"
#include "pps/gpu_cfr.h"
#include "pps/postflop_cfr.h"
#include "pps/range.h"
#include "pps/server.h"
#include "solver/solution_file.h"

// A solver of its own, or a spot served by a SpotServer (then queries run
// between the server's steps and solves are batched jobs of the server).
struct pps_solver {
  std::unique_ptr<pps::gpu::GpuPostflopSolver> s;
  std::shared_ptr<pps::gpu::SpotServer> srv;
  int id = -1;
  ~pps_solver() {
    if (srv) {
      try {
        srv->remove(id);
      } catch (...) {
      }
    }
  }
};

struct pps_server {
  std::shared_ptr<pps::gpu::SpotServer> srv;
};

namespace {

thread_local std::string g_err;

template <typename F>
int guard(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
  } catch (...) {
    g_err = "unknown error";
  }
  return -1;
}

pps::gpu::PostflopSpot toSpot(const pps_spot* s) {
  if (!s) throw std::runtime_error("null spot");
  if (s->n_board < 3 || s->n_board > 5) throw std::runtime_error("board must be 3, 4 or 5 cards");
  pps::gpu::PostflopSpot out;
  out.nBoard = s->n_board;
  for (int i = 0; i < s->n_board; ++i) {
    if (s->board[i] >= 52) throw std::runtime_error("bad card id");
    out.board[i] = s->board[i];
  }
  for (int i = 0; i < pps::kCombos; ++i) {
    out.oop.w[i] = s->oop[i];
    out.ip.w[i] = s->ip[i];
  }
  out.pot = s->pot;
  out.stack = s->stack;
  return out;
}

pps::pf::BetConfig toCfg(const pps_bet_config* c) {
  if (!c) throw std::runtime_error("null bet config");
  if (c->n_bets < 0 || c->n_bets > 16 || c->n_raises < 0 || c->n_raises > 16)
    throw std::runtime_error("bad bet config sizes");
  pps::pf::BetConfig cfg;
  cfg.betFracs.assign(c->bets, c->bets + c->n_bets);
  cfg.raiseMults.assign(c->raises, c->raises + c->n_raises);
  cfg.betAllIn = c->bet_allin != 0;
  cfg.betGeometric = c->bet_geo != 0;
  cfg.maxRaises = c->max_raises;
  cfg.addAllinThreshold = c->add_allin;
  cfg.forceAllinThreshold = c->force_allin;
  cfg.mergingThreshold = c->merging;
  return cfg;
}

// Runs f(solver): directly, or on the server between its steps.
template <typename F>
auto on(pps_solver* h, F&& f) -> decltype(f(std::declval<pps::gpu::GpuPostflopSolver&>())) {
  using R = decltype(f(std::declval<pps::gpu::GpuPostflopSolver&>()));
  if (!h || (!h->s && !h->srv)) throw std::runtime_error("null solver");
  if (!h->srv) return f(*h->s);
  std::optional<R> r;
  h->srv->with(h->id, [&](pps::gpu::GpuPostflopSolver& g) { r.emplace(f(g)); });
  return std::move(*r);
}

pps_solver* served(pps_solver* h, const char* what) {
  if (!h || !h->srv) throw std::runtime_error(std::string(what) + ": not a served solver");
  return h;
}

// Blocks until the served spot's job ends; a job error becomes the call's.
void finishJob(pps_solver* h) {
  const pps::gpu::JobInfo j = h->srv->wait(h->id);
  if (!j.error.empty()) throw std::runtime_error(j.error);
}

}  // namespace

extern "C" {

const char* pps_last_error(void) { return g_err.c_str(); }

uint64_t pps_hand_key(const uint8_t* cards, int n) {
  if (!cards || n < 5 || n > 7) return 0;
  return pps::handKeyN(cards, n);
}

int pps_parse_range(const char* text, double* w) {
  return guard([&] {
    if (!text || !w) throw std::runtime_error("null argument");
    const pps::Range r = pps::parseRange(text);
    std::memcpy(w, r.w, sizeof(r.w));
    return 0;
  });
}

int pps_bet_actions(int64_t pot_base, int64_t stack, const pps_bet_config* cfg, int64_t sc0,
                    int64_t sc1, int actor, int after_allin, int num_bets, int num_streets,
                    int64_t* out, int cap) {
  return guard([&] {
    const auto acts = pps::pf::betActions(pot_base, stack, toCfg(cfg), sc0, sc1, actor,
                                          after_allin != 0, num_bets, num_streets);
    if ((int)acts.size() * 2 > cap) throw std::runtime_error("output too small");
    for (size_t i = 0; i < acts.size(); ++i) {
      out[2 * i] = (int64_t)acts[i].kind;
      out[2 * i + 1] = acts[i].amount;
    }
    return (int)acts.size();
  });
}

int pps_schedule(const char* algo, int t, int iters, double out3[3]) {
  return guard([&] {
    const std::string a = algo ? algo : "";
    if (a != "dcfr" && a != "hs30" && a != "dcfr-g2") throw std::runtime_error("unknown algo");
    pps::gpu::discountCoefs(t, iters, a, out3);
    return 0;
  });
}

int pps_compile_info(const pps_spot* spot, const pps_bet_config* cfg, int64_t out[8],
                     double* bytes) {
  return pps_compile_info_ex(spot, cfg, 0, out, bytes);
}

int pps_compile_info_ex(const pps_spot* spot, const pps_bet_config* cfg, int flags,
                        int64_t out[8], double* bytes) {
  return guard([&] {
    if (flags & ~PPS_SUIT_ISOMORPHISM) throw std::runtime_error("unknown option flags");
    const pps::gpu::CompiledTree t =
        pps::gpu::compileSpot(toSpot(spot), toCfg(cfg), (flags & PPS_SUIT_ISOMORPHISM) != 0);
    if (t.empty) throw std::runtime_error("empty ranges");
    int64_t sd = 0, fd = 0, ch = 0;
    for (uint8_t k : t.kind) {
      sd += k == PPSG_SHOWDOWN;
      fd += k == PPSG_FOLD;
      ch += k == PPSG_CHANCE;
    }
    out[0] = t.numNodes;
    out[1] = t.maxDepth;
    out[2] = (int64_t)t.decideNodes.size();
    out[3] = t.regTotal;
    out[4] = sd;
    out[5] = fd;
    out[6] = ch;
    out[7] = (int64_t)t.clsNBoard.size();
    if (bytes) *bytes = t.iterationBytes();
    return 0;
  });
}

int pps_solver_create(const pps_spot* spot, const pps_bet_config* cfg, pps_solver** out) {
  return guard([&] {
    if (!out) throw std::runtime_error("null output");
    auto h = std::make_unique<pps_solver>();
    h->s = std::make_unique<pps::gpu::GpuPostflopSolver>(toSpot(spot), toCfg(cfg));
    *out = h.release();
    return 0;
  });
}

int pps_solver_create_ex(const pps_spot* spot, const pps_bet_config* cfg, int flags,
                         pps_solver** out) {
  return guard([&] {
    if (!out) throw std::runtime_error("null output");
    if (flags & ~PPS_SUIT_ISOMORPHISM) throw std::runtime_error("unknown option flags");
    pps::gpu::SolverOptions opt;
    opt.suitIsomorphism = (flags & PPS_SUIT_ISOMORPHISM) != 0;
    auto h = std::make_unique<pps_solver>();
    h->s = std::make_unique<pps::gpu::GpuPostflopSolver>(toSpot(spot), toCfg(cfg), opt);
    *out = h.release();
    return 0;
  });
}

int pps_solver_load(const char* path, pps_solver** out) {
  return guard([&] {
    if (!out || !path) throw std::runtime_error("null argument");
    auto h = std::make_unique<pps_solver>();
    h->s = std::make_unique<pps::gpu::GpuPostflopSolver>(std::string(path));
    *out = h.release();
    return 0;
  });
}

void pps_solver_destroy(pps_solver* s) { delete s; }

int pps_solver_solve(pps_solver* s, int iters, const char* algo, double target) {
  return guard([&] {
    const std::string a = algo ? algo : "dcfr";
    if (s && s->srv) {
      s->srv->solve(s->id, iters, a, target);
      finishJob(s);
      return 0;
    }
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      g.solve(iters, a, target);
      return 0;
    });
  });
}

int pps_solver_continue(pps_solver* s, int iters, double target) {
  return guard([&] {
    if (s && s->srv) {
      s->srv->continueSolve(s->id, iters, target);
      finishJob(s);
      return 0;
    }
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      g.continueSolve(iters, target);
      return 0;
    });
  });
}

int pps_solver_reset(pps_solver* s) {
  return guard([&] {
    if (s && s->srv) {
      s->srv->reset(s->id);
      return 0;
    }
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      g.reset();
      return 0;
    });
  });
}

int pps_solver_iterations_run(pps_solver* s) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) { return g.iterationsRun(); });
  });
}

int64_t pps_solver_total_iterations(pps_solver* s) {
  int64_t v = -1;
  guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      v = g.totalIterations();
      return 0;
    });
  });
  return v;
}

int pps_solver_stats(pps_solver* s, double out4[4]) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      const pps::pf::NodeStats st = g.stats();
      out4[0] = st.ev0;
      out4[1] = st.ev1;
      out4[2] = st.expl;
      out4[3] = (double)st.pairMass;
      return 0;
    });
  });
}

static int copyOut(const std::vector<double>& v, double* out, int cap) {
  if ((int)v.size() > cap) throw std::runtime_error("output too small");
  std::copy(v.begin(), v.end(), out);
  return (int)v.size();
}

int pps_solver_root_strategy(pps_solver* s, double* out, int cap) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) { return copyOut(g.rootStrategy(), out, cap); });
  });
}

int pps_solver_node_strategy(pps_solver* s, int idx, double* out, int cap) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) { return copyOut(g.nodeStrategy(idx), out, cap); });
  });
}

int pps_solver_num_decide_nodes(pps_solver* s) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) { return g.numDecideNodes(); });
  });
}

int pps_solver_decide_node(pps_solver* s, int idx, int32_t info6[6], int64_t* actions, int cap) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      const pps::gpu::DecideNodeInfo d = g.decideNode(idx);
      info6[0] = d.nodeId;
      info6[1] = d.player;
      info6[2] = d.depth;
      info6[3] = d.nBoard;
      info6[4] = d.nCombos;
      info6[5] = (int32_t)d.actions.size();
      if ((int)d.actions.size() * 3 > cap) throw std::runtime_error("output too small");
      for (size_t a = 0; a < d.actions.size(); ++a) {
        actions[3 * a] = (int64_t)d.actions[a].kind;
        actions[3 * a + 1] = d.actions[a].amount;
        actions[3 * a + 2] = d.children[a];
      }
      return 0;
    });
  });
}

int pps_solver_combo_strategy(pps_solver* s, int idx, int32_t* combos, double* freqs, int cap) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      const pps::gpu::ComboStrategy cs = g.comboStrategy(idx);
      if (cs.combos.size() > (size_t)cap || cs.freqs.size() > (size_t)cap * 32)
        throw std::runtime_error("output too small");
      std::copy(cs.combos.begin(), cs.combos.end(), combos);
      std::copy(cs.freqs.begin(), cs.freqs.end(), freqs);
      return (int)cs.combos.size();
    });
  });
}

int pps_solver_num_nodes(pps_solver* s) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) { return g.numNodes(); });
  });
}

int pps_solver_max_depth(pps_solver* s) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) { return g.maxDepth(); });
  });
}

int pps_solver_num_base_combos(pps_solver* s, int player) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) { return g.numBaseCombos(player); });
  });
}

int pps_solver_save(pps_solver* s, const char* path) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      g.save(path ? path : "");
      return 0;
    });
  });
}

int64_t pps_solver_read_rows(pps_solver* s, int which, float* out, int64_t cap) {
  int64_t n = -1;
  guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      const std::vector<float> r = g.rows(which);
      if (!out) {  // size query
        n = (int64_t)r.size();
        return 0;
      }
      if ((int64_t)r.size() > cap) throw std::runtime_error("output too small");
      std::copy(r.begin(), r.end(), out);
      n = (int64_t)r.size();
      return 0;
    });
  });
  return n;
}

int pps_solver_write_rows(pps_solver* s, int which, const float* in, int64_t n) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      g.setRows(which, std::vector<float>(in, in + n));
      return 0;
    });
  });
}

int pps_solver_half_step(pps_solver* s, int tr, int t, const char* algo, int planned) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      g.halfStep(tr, t, algo ? algo : "dcfr", planned);
      return 0;
    });
  });
}

int pps_solver_root_value(pps_solver* s, int tr, int mode, float* root_cfv, double* value) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      std::vector<float> cfv;
      const double v = g.rootValue(tr, mode, &cfv);
      if (root_cfv) std::copy(cfv.begin(), cfv.end(), root_cfv);
      if (value) *value = v;
      return 0;
    });
  });
}

double pps_solver_iteration_bytes(pps_solver* s) {
  double v = -1.0;
  guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      v = g.iterationBytes();
      return 0;
    });
  });
  return v;
}

int pps_solver_profile(pps_solver* s, int iters, int detail, double* out) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      const std::vector<double> v = g.profileIterations(iters, detail != 0);
      for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
      return 0;
    });
  });
}

int pps_solver_engine_info(pps_solver* s, int64_t out6[6]) {
  return guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      const std::vector<int64_t> v = g.engineInfo();
      for (size_t i = 0; i < 6; ++i) out6[i] = v[i];
      return 0;
    });
  });
}

int pps_solver_engine(pps_solver* s, ppsg_engine** engine, int* slot) {
  return guard([&] {
    if (!engine || !slot) throw std::runtime_error("null output");
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      ppsg_engine* e = g.engineHandle();
      if (!e) throw std::runtime_error("solver compiled to empty ranges (no engine)");
      *engine = e;
      *slot = g.engineSlot();
      return 0;
    });
  });
}

void pps_serve_defaults(pps_serve_options* opt) {
  if (!opt) return;
  const pps::gpu::ServeOptions d;
  opt->target_frac = d.targetFrac;
  opt->check_every = d.checkEvery;
  opt->max_iters = d.maxIters;
  opt->arena_bytes = d.arenaBytes;
  opt->max_spots = d.maxSpots;
  opt->max_live = d.maxLive;
  opt->compile_threads = d.compileThreads;
  opt->real_time = d.realTime ? 1 : 0;
  opt->prefill = d.prefill ? 1 : 0;
  opt->flags = 0;
  opt->shapes = nullptr;
  opt->n_shapes = 0;
  opt->term_split = 0;
  opt->request_flags = nullptr;
}

int pps_serve(int n, const pps_spot* spots, const pps_bet_config* cfgs, const double* arrival_ms,
              const char* algo, const pps_serve_options* opt, pps_serve_result* out,
              pps_serve_report* report) {
  return guard([&] {
    if (n < 0 || (n > 0 && (!spots || !cfgs || !out))) throw std::runtime_error("null argument");
    pps::gpu::ServeOptions o;
    pps::gpu::SolverOptions spotOpt;
    if (opt) {
      if (opt->flags & ~PPS_SUIT_ISOMORPHISM) throw std::runtime_error("unknown option flags");
      spotOpt.suitIsomorphism = (opt->flags & PPS_SUIT_ISOMORPHISM) != 0;
      o.targetFrac = opt->target_frac;
      o.checkEvery = opt->check_every;
      o.maxIters = opt->max_iters;
      o.arenaBytes = opt->arena_bytes;
      o.maxSpots = opt->max_spots;
      o.maxLive = opt->max_live;
      o.compileThreads = opt->compile_threads;
      o.realTime = opt->real_time != 0;
      o.prefill = opt->prefill != 0;
      if (opt->shapes && opt->n_shapes > 0) o.shapes.assign(opt->shapes, opt->shapes + opt->n_shapes);
      o.termSplit = opt->term_split;
    }
    if (algo) o.algo = algo;
    std::vector<pps::gpu::SpotRequest> reqs(n);
    const int32_t* perReq = opt ? opt->request_flags : nullptr;
    for (int i = 0; i < n; ++i) {
      reqs[i].spot = toSpot(spots + i);
      reqs[i].cfg = toCfg(cfgs + i);
      reqs[i].arrivalMs = arrival_ms ? arrival_ms[i] : 0.0;
      reqs[i].options = spotOpt;
      if (perReq) {
        if (perReq[i] & ~PPS_SUIT_ISOMORPHISM)
          throw std::runtime_error("request " + std::to_string(i) + ": unknown option flags");
        reqs[i].options.suitIsomorphism = (perReq[i] & PPS_SUIT_ISOMORPHISM) != 0;
      }
    }
    pps::gpu::ServeReport rep;
    const std::vector<pps::gpu::SpotResult> res = pps::gpu::serve(reqs, o, &rep);
    if (report) std::memset(report, 0, sizeof(*report));
    for (int i = 0; i < n; ++i) {
      const pps::gpu::SpotResult& r = res[i];
      pps_serve_result& d = out[i];
      std::memset(&d, 0, sizeof(d));
      d.status = r.ok ? 0 : -1;
      d.converged = r.converged ? 1 : 0;
      d.iterations = r.iterations;
      d.n_root_actions = (int32_t)std::min<size_t>(r.rootStrategy.size(), 32);
      d.expl = r.expl;
      d.ev0 = r.ev0;
      d.ev1 = r.ev1;
      for (int a = 0; a < d.n_root_actions; ++a) d.root_strategy[a] = r.rootStrategy[a];
      d.nodes = r.numNodes;
      d.footprint = r.footprint;
      d.arrival_ms = r.arrivalMs;
      d.admit_ms = r.admitMs;
      d.retire_ms = r.retireMs;
      std::strncpy(d.error, r.error.c_str(), sizeof(d.error) - 1);
    }
    if (report) {
      report->wall_ms = rep.wallMs;
      report->busy_ms = rep.busyMs;
      report->device_ms = rep.deviceMs;
      report->walk_ms = rep.walkMs;
      report->admit_ms = rep.admitMs;
      report->compact_ms = rep.compactMs;
      report->iterations = rep.iterations;
      report->steps = rep.steps;
      report->peak_live = rep.peakLive;
      report->compactions = rep.compactions;
      report->group_size = rep.groupSize;
      for (size_t k = 0; k < rep.stepLive.size() && k < PPS_SERVE_TRACE; ++k) {
        report->step_live[k] = rep.stepLive[k];
        report->step_device_ms[k] = rep.stepDeviceMs[k];
        report->step_check_ms[k] = rep.stepCheckMs[k];
      }
    }
    return 0;
  });
}

int pps_batch_profile(int n, const pps_spot* spots, const pps_bet_config* cfgs, const char* algo,
                      int iters, int warmup, int64_t arena_bytes, int detail, double out[42],
                      int32_t* shapes, int32_t shapes_cap) {
  return guard([&] {
    if (n <= 0 || !spots || !cfgs || !out) throw std::runtime_error("null argument");
    std::vector<pps::gpu::SpotRequest> reqs(n);
    for (int i = 0; i < n; ++i) {
      reqs[i].spot = toSpot(spots + i);
      reqs[i].cfg = toCfg(cfgs + i);
    }
    const pps::gpu::BatchProfile P =
        pps::gpu::profileBatch(reqs, algo ? algo : "dcfr", iters, arena_bytes, detail != 0,
                               warmup);
    const double v[16] = {(double)P.spots, (double)P.iters, P.graphMsPerIter, P.fwdMs,
                          P.fwdLaunches, P.bwdMs, P.bwdLaunches, P.fwdBytes, P.bwdBytes,
                          (double)P.kernelsPerIter, (double)P.itemsPerIter, (double)P.arenaBytes,
                          P.decideMs, P.chanceMs, P.foldMs, P.showdownMs};
    for (int i = 0; i < 16; ++i) out[i] = v[i];
    out[16] = P.groupSize;
    for (int i = 0; i < pps::gpu::kNumGroupCandidates; ++i) out[17 + i] = P.tuneMs[i];
    out[23] = P.fwdDecideMs;
    out[24] = P.fwdChanceMs;
    for (int k = 0; k < 6; ++k) out[25 + k] = P.kindBytes[k];
    out[31] = P.foldLaunches;
    out[32] = P.showdownLaunches;
    out[33] = P.termSplit;
    for (int i = 0; i < pps::gpu::kNumSplitCandidates; ++i) out[34 + i] = P.splitMs[i];
    if (shapes) {
      if ((size_t)shapes_cap < P.shapes.size())
        throw std::runtime_error("profile: the shape table needs " + std::to_string(P.shapes.size()) +
                                 " entries");
      std::copy(P.shapes.begin(), P.shapes.end(), shapes);
    }
    return 0;
  });
}

int pps_solution_inspect(const char* path, pps_solution_info* out) {
  return guard([&] {
    if (!path || !out) throw std::runtime_error("null argument");
    const pps::gpu::Solution s = pps::gpu::readSolution(path, false);
    std::memset(out, 0, sizeof(*out));
    out->n_board = s.spot.nBoard;
    for (int i = 0; i < 5; ++i) out->board[i] = s.spot.board[i];
    out->pot = s.spot.pot;
    out->stack = s.spot.stack;
    out->n_bets = (int32_t)s.cfg.betFracs.size();
    out->n_raises = (int32_t)s.cfg.raiseMults.size();
    out->nodes = s.nodes;
    out->iterations = s.iterations;
    out->rows = s.rows;
    std::strncpy(out->algo, s.algo.c_str(), sizeof(out->algo) - 1);
    return 0;
  });
}

int pps_compile_footprint(const pps_spot* spot, const pps_bet_config* cfg, int64_t* bytes) {
  return guard([&] {
    if (!bytes) throw std::runtime_error("null argument");
    const pps::gpu::CompiledTree t = pps::gpu::compileSpot(toSpot(spot), toCfg(cfg));
    if (t.empty) throw std::runtime_error("empty ranges");
    const ppsg_tree_desc d = t.desc();
    if (ppsg_footprint(&d, bytes) != 0) throw std::runtime_error(ppsg_last_error());
    return 0;
  });
}

int pps_server_create(const pps_server_options* opt, pps_server** out) {
  return guard([&] {
    if (!out) throw std::runtime_error("null output");
    pps::gpu::ServerOptions o;
    if (opt) {
      o.arenaBytes = opt->arena_bytes;
      if (opt->max_spots > 0) o.maxSpots = opt->max_spots;
      if (opt->max_step > 0) o.maxStep = opt->max_step;
    }
    auto h = std::make_unique<pps_server>();
    h->srv = std::make_shared<pps::gpu::SpotServer>(o);
    *out = h.release();
    return 0;
  });
}

void pps_server_destroy(pps_server* srv) { delete srv; }

static int admitted(pps_server* srv, pps_solver** out, const std::function<int()>& add) {
  if (!srv || !srv->srv || !out) throw std::runtime_error("null argument");
  int id = -1;
  try {
    id = add();
  } catch (const pps::gpu::ArenaFull& e) {
    g_err = e.what();
    return PPS_ARENA_FULL;
  }
  auto h = std::make_unique<pps_solver>();
  h->srv = srv->srv;
  h->id = id;
  *out = h.release();
  return 0;
}

int pps_server_add(pps_server* srv, const pps_spot* spot, const pps_bet_config* cfg,
                   pps_solver** out) {
  return guard([&] {
    const pps::gpu::PostflopSpot sp = toSpot(spot);
    const pps::pf::BetConfig c = toCfg(cfg);
    return admitted(srv, out, [&] { return srv->srv->add(sp, c); });
  });
}

int pps_server_add_ex(pps_server* srv, const pps_spot* spot, const pps_bet_config* cfg, int flags,
                      pps_solver** out) {
  return guard([&] {
    if (flags & ~PPS_SUIT_ISOMORPHISM) throw std::runtime_error("unknown option flags");
    const pps::gpu::PostflopSpot sp = toSpot(spot);
    const pps::pf::BetConfig c = toCfg(cfg);
    pps::gpu::SolverOptions opt;
    opt.suitIsomorphism = (flags & PPS_SUIT_ISOMORPHISM) != 0;
    return admitted(srv, out, [&] { return srv->srv->add(sp, c, opt); });
  });
}

int pps_server_load(pps_server* srv, const char* path, pps_solver** out) {
  return guard([&] {
    if (!path) throw std::runtime_error("null path");
    const std::string p(path);
    return admitted(srv, out, [&] { return srv->srv->load(p); });
  });
}

int pps_server_stats(pps_server* srv, pps_server_counters* out) {
  return guard([&] {
    if (!srv || !srv->srv || !out) throw std::runtime_error("null argument");
    const pps::gpu::ServerStats s = srv->srv->stats();
    std::memset(out, 0, sizeof(*out));
    out->steps = s.steps;
    out->iterations = s.iterations;
    out->device_ms = s.deviceMs;
    out->check_ms = s.checkMs;
    out->resident = s.resident;
    out->running = s.running;
    out->peak_running = s.peakRunning;
    out->arena_bytes = s.arenaBytes;
    out->free_bytes = s.freeBytes;
    return 0;
  });
}

int pps_solver_start_solve(pps_solver* s, int iters, const char* algo, double target) {
  return guard([&] {
    served(s, "start_solve")->srv->solve(s->id, iters, algo ? algo : "dcfr", target);
    return 0;
  });
}

int pps_solver_start_continue(pps_solver* s, int iters, double target) {
  return guard([&] {
    served(s, "start_continue")->srv->continueSolve(s->id, iters, target);
    return 0;
  });
}

int pps_solver_job(pps_solver* s, int wait, pps_job* out) {
  return guard([&] {
    if (!out) throw std::runtime_error("null output");
    served(s, "job");
    const pps::gpu::JobInfo j = wait ? s->srv->wait(s->id) : s->srv->job(s->id);
    std::memset(out, 0, sizeof(*out));
    out->running = j.running ? 1 : 0;
    out->status = j.status == "running"     ? PPS_JOB_RUNNING
                  : j.status == "done"      ? PPS_JOB_DONE
                  : j.status == "cancelled" ? PPS_JOB_CANCELLED
                  : j.status == "error"     ? PPS_JOB_ERROR
                                            : PPS_JOB_NONE;
    out->kind = j.kind == "solve" ? 1 : j.kind == "continue" ? 2 : 0;
    out->start = j.start;
    out->done = j.done;
    out->total = j.total;
    out->target = j.target;
    if (!j.error.empty()) throw std::runtime_error(j.error);
    return 0;
  });
}

int pps_solver_cancel(pps_solver* s) {
  return guard([&] { return served(s, "cancel")->srv->cancel(s->id) ? 1 : 0; });
}

double pps_solver_last_solve_ms(pps_solver* s) {
  double v = -1.0;
  guard([&] {
    return on(s, [&](pps::gpu::GpuPostflopSolver& g) {
      v = g.lastSolveDeviceMs();
      return 0;
    });
  });
  return v;
}

}  // extern "C"

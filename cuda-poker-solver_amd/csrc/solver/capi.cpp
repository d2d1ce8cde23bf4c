// C ABI over GpuPostflopSolver and the integer layer (include/pps_capi.h).
// Exceptions never cross the boundary: every entry returns a status and
// leaves the message in a thread-local buffer.
#include "pps_capi.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "host/compile.h"
#include "pps/batch.h"
#include "pps/cards.h"
#include "pps/gpu_cfr.h"
#include "pps/postflop_cfr.h"
#include "pps/range.h"

struct pps_solver {
  std::unique_ptr<pps::gpu::GpuPostflopSolver> s;
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

pps::gpu::GpuPostflopSolver& S(pps_solver* h) {
  if (!h || !h->s) throw std::runtime_error("null solver");
  return *h->s;
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
  return guard([&] {
    const pps::gpu::CompiledTree t = pps::gpu::compileSpot(toSpot(spot), toCfg(cfg));
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
    S(s).solve(iters, algo ? algo : "dcfr", target);
    return 0;
  });
}

int pps_solver_continue(pps_solver* s, int iters, double target) {
  return guard([&] {
    S(s).continueSolve(iters, target);
    return 0;
  });
}

int pps_solver_reset(pps_solver* s) {
  return guard([&] {
    S(s).reset();
    return 0;
  });
}

int pps_solver_iterations_run(pps_solver* s) {
  return guard([&] { return S(s).iterationsRun(); });
}

int64_t pps_solver_total_iterations(pps_solver* s) {
  int64_t v = -1;
  guard([&] {
    v = S(s).totalIterations();
    return 0;
  });
  return v;
}

int pps_solver_stats(pps_solver* s, double out4[4]) {
  return guard([&] {
    const pps::pf::NodeStats st = S(s).stats();
    out4[0] = st.ev0;
    out4[1] = st.ev1;
    out4[2] = st.expl;
    out4[3] = (double)st.pairMass;
    return 0;
  });
}

static int copyOut(const std::vector<double>& v, double* out, int cap) {
  if ((int)v.size() > cap) throw std::runtime_error("output too small");
  std::copy(v.begin(), v.end(), out);
  return (int)v.size();
}

int pps_solver_root_strategy(pps_solver* s, double* out, int cap) {
  return guard([&] { return copyOut(S(s).rootStrategy(), out, cap); });
}

int pps_solver_node_strategy(pps_solver* s, int idx, double* out, int cap) {
  return guard([&] { return copyOut(S(s).nodeStrategy(idx), out, cap); });
}

int pps_solver_num_decide_nodes(pps_solver* s) {
  return guard([&] { return S(s).numDecideNodes(); });
}

int pps_solver_decide_node(pps_solver* s, int idx, int32_t info6[6], int64_t* actions, int cap) {
  return guard([&] {
    const pps::gpu::DecideNodeInfo d = S(s).decideNode(idx);
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
}

int pps_solver_combo_strategy(pps_solver* s, int idx, int32_t* combos, double* freqs, int cap) {
  return guard([&] {
    const pps::gpu::ComboStrategy cs = S(s).comboStrategy(idx);
    if (cs.combos.size() > (size_t)cap || cs.freqs.size() > (size_t)cap * 32)
      throw std::runtime_error("output too small");
    std::copy(cs.combos.begin(), cs.combos.end(), combos);
    std::copy(cs.freqs.begin(), cs.freqs.end(), freqs);
    return (int)cs.combos.size();
  });
}

int pps_solver_num_nodes(pps_solver* s) {
  return guard([&] { return S(s).numNodes(); });
}

int pps_solver_max_depth(pps_solver* s) {
  return guard([&] { return S(s).maxDepth(); });
}

int pps_solver_num_base_combos(pps_solver* s, int player) {
  return guard([&] { return S(s).numBaseCombos(player); });
}

int pps_solver_save(pps_solver* s, const char* path) {
  return guard([&] {
    S(s).save(path ? path : "");
    return 0;
  });
}

int64_t pps_solver_read_rows(pps_solver* s, int which, float* out, int64_t cap) {
  int64_t n = -1;
  guard([&] {
    const std::vector<float> r = S(s).rows(which);
    if (!out) {  // size query
      n = (int64_t)r.size();
      return 0;
    }
    if ((int64_t)r.size() > cap) throw std::runtime_error("output too small");
    std::copy(r.begin(), r.end(), out);
    n = (int64_t)r.size();
    return 0;
  });
  return n;
}

int pps_solver_write_rows(pps_solver* s, int which, const float* in, int64_t n) {
  return guard([&] {
    S(s).setRows(which, std::vector<float>(in, in + n));
    return 0;
  });
}

int pps_solver_half_step(pps_solver* s, int tr, int t, const char* algo, int planned) {
  return guard([&] {
    S(s).halfStep(tr, t, algo ? algo : "dcfr", planned);
    return 0;
  });
}

int pps_solver_root_value(pps_solver* s, int tr, int mode, float* root_cfv, double* value) {
  return guard([&] {
    std::vector<float> cfv;
    const double v = S(s).rootValue(tr, mode, &cfv);
    if (root_cfv) std::copy(cfv.begin(), cfv.end(), root_cfv);
    if (value) *value = v;
    return 0;
  });
}

double pps_solver_iteration_bytes(pps_solver* s) {
  double v = -1.0;
  guard([&] {
    v = S(s).iterationBytes();
    return 0;
  });
  return v;
}

int pps_solver_profile(pps_solver* s, int iters, int detail, double* out) {
  return guard([&] {
    const std::vector<double> v = S(s).profileIterations(iters, detail != 0);
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
    return 0;
  });
}

int pps_solver_engine_info(pps_solver* s, int64_t out6[6]) {
  return guard([&] {
    const std::vector<int64_t> v = S(s).engineInfo();
    for (size_t i = 0; i < 6; ++i) out6[i] = v[i];
    return 0;
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
}

int pps_serve(int n, const pps_spot* spots, const pps_bet_config* cfgs, const double* arrival_ms,
              const char* algo, const pps_serve_options* opt, pps_serve_result* out,
              pps_serve_report* report) {
  return guard([&] {
    if (n < 0 || (n > 0 && (!spots || !cfgs || !out))) throw std::runtime_error("null argument");
    pps::gpu::ServeOptions o;
    if (opt) {
      o.targetFrac = opt->target_frac;
      o.checkEvery = opt->check_every;
      o.maxIters = opt->max_iters;
      o.arenaBytes = opt->arena_bytes;
      o.maxSpots = opt->max_spots;
      o.maxLive = opt->max_live;
      o.compileThreads = opt->compile_threads;
      o.realTime = opt->real_time != 0;
      o.prefill = opt->prefill != 0;
    }
    if (algo) o.algo = algo;
    std::vector<pps::gpu::SpotRequest> reqs(n);
    for (int i = 0; i < n; ++i) {
      reqs[i].spot = toSpot(spots + i);
      reqs[i].cfg = toCfg(cfgs + i);
      reqs[i].arrivalMs = arrival_ms ? arrival_ms[i] : 0.0;
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
                      int iters, int64_t arena_bytes, int detail, double out[33]) {
  return guard([&] {
    if (n <= 0 || !spots || !cfgs || !out) throw std::runtime_error("null argument");
    std::vector<pps::gpu::SpotRequest> reqs(n);
    for (int i = 0; i < n; ++i) {
      reqs[i].spot = toSpot(spots + i);
      reqs[i].cfg = toCfg(cfgs + i);
    }
    const pps::gpu::BatchProfile P =
        pps::gpu::profileBatch(reqs, algo ? algo : "dcfr", iters, arena_bytes, detail != 0);
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

double pps_solver_last_solve_ms(pps_solver* s) {
  double v = -1.0;
  guard([&] {
    v = S(s).lastSolveDeviceMs();
    return 0;
  });
  return v;
}

}  // extern "C"

// pps_native: Python binding of pps::gpu::GpuPostflopSolver — drop-in for
// the reference's python/pps_native.cpp:514-550 (same class, methods,
// argument names and defaults, same result dicts) — plus pps.Server, which
// serves many spots on one GPU with their solves batched
// (include/pps/server.h). Engine calls release the GIL; callers serialize
// use of one Solver.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <memory>
#include <optional>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <vector>

#include "pps/gpu_cfr.h"
#include "pps/postflop_cfr.h"
#include "pps/range.h"
#include "pps/server.h"

namespace py = pybind11;
using pps::Card;
using pps::Range;
using pps::gpu::GpuPostflopSolver;
using pps::gpu::SpotServer;

namespace {

std::string cardText(int c) {
  static const char kR[] = "23456789TJQKA";
  static const char kS[] = "cdhs";
  return std::string{kR[c % 13], kS[c / 13]};
}

Card readCard(const std::string& s, size_t i) {
  if (i + 1 >= s.size() || s[i] == ',') throw std::runtime_error("bad card text '" + s.substr(0, 2) + "'");
  const std::string ranks = "23456789TJQKA", suits = "cdhs";
  const size_t r = ranks.find((char)std::toupper((unsigned char)s[i]));
  if (r == std::string::npos) throw std::runtime_error("bad card rank '" + std::string(1, s[i]) + "'");
  const size_t u = suits.find((char)std::tolower((unsigned char)s[i + 1]));
  if (u == std::string::npos)
    throw std::runtime_error("bad card suit '" + std::string(1, s[i + 1]) + "'");
  return (Card)(u * 13 + r);
}

bool listLike(const py::object& o) { return py::isinstance<py::list>(o) || py::isinstance<py::tuple>(o); }

std::vector<Card> readBoard(const py::object& board) {
  std::string text;
  if (py::isinstance<py::str>(board)) {
    text = board.cast<std::string>();
  } else if (listLike(board)) {
    for (auto c : board) text += py::str(c).cast<std::string>();
  } else {
    throw std::runtime_error("board must be a string or a list of card strings");
  }
  if (text.size() % 2 != 0 || text.size() < 6 || text.size() > 10)
    throw std::runtime_error("board must be 3, 4 or 5 cards (e.g. \"Qs9h2d\")");
  std::vector<Card> out;
  for (size_t i = 0; i < text.size(); i += 2) out.push_back(readCard(text, i));
  return out;
}

Range readRange(const py::object& r) {
  if (py::isinstance<py::str>(r)) return pps::parseRange(r.cast<std::string>());
  if (py::isinstance<py::dict>(r)) {
    std::string text;
    for (auto kv : r.cast<py::dict>()) {
      if (!text.empty()) text += ",";
      text += py::str(kv.first).cast<std::string>() + ":" + std::to_string(kv.second.cast<double>());
    }
    return pps::parseRange(text);
  }
  if (listLike(r)) {
    const auto v = r.cast<std::vector<double>>();
    if (v.size() != (size_t)pps::kCombos) throw std::runtime_error("range list must have exactly 1326 weights");
    Range out;
    for (int i = 0; i < pps::kCombos; ++i) out.w[i] = v[i];
    return out;
  }
  throw std::runtime_error("range must be a string, dict or 1326-float list");
}

void readBetToken(const std::string& t, pps::pf::BetConfig& cfg) {
  if (t == "a" || t == "A" || t == "allin") cfg.betAllIn = true;
  else if (t == "e" || t == "E" || t == "geo") cfg.betGeometric = true;
  else {
    try {
      cfg.betFracs.push_back(std::stod(t));
    } catch (...) {
      throw std::runtime_error("bad bet size token '" + t + "'");
    }
  }
}

std::vector<std::string> splitComma(const std::string& s) {
  std::vector<std::string> out;
  size_t pos = 0;
  while (true) {
    const size_t c = s.find(',', pos);
    std::string t = c == std::string::npos ? s.substr(pos) : s.substr(pos, c - pos);
    if (!t.empty()) out.push_back(t);
    if (c == std::string::npos) break;
    pos = c + 1;
  }
  return out;
}

void readBets(const py::object& arg, pps::pf::BetConfig& cfg) {
  if (py::isinstance<py::str>(arg)) {
    for (const std::string& t : splitComma(arg.cast<std::string>())) readBetToken(t, cfg);
  } else if (listLike(arg)) {
    for (auto item : arg) {
      if (py::isinstance<py::str>(item)) {
        const std::string t = item.cast<std::string>();
        if (t == "a" || t == "A" || t == "allin" || t == "e" || t == "E" || t == "geo")
          readBetToken(t, cfg);
        else
          throw std::runtime_error("bad bet size token '" + t + "'");
      } else {
        cfg.betFracs.push_back(item.cast<double>());
      }
    }
  } else {
    throw std::runtime_error("bets must be a string or a list");
  }
}

std::vector<double> readDoubles(const py::object& arg) {
  if (py::isinstance<py::str>(arg)) {
    std::vector<double> out;
    for (const std::string& t : splitComma(arg.cast<std::string>())) out.push_back(std::stod(t));
    return out;
  }
  return arg.cast<std::vector<double>>();
}

py::dict actionDict(const pps::gpu::ActionLabel& a) {
  static const char* kNames[] = {"fold", "check", "call", "bet", "raise", "allin"};
  py::dict d;
  d["kind"] = kNames[(int)a.kind];
  d["amount"] = a.amount;
  return d;
}

py::array_t<double> arr(const std::vector<double>& v) {
  py::array_t<double> a((py::ssize_t)v.size());
  std::copy(v.begin(), v.end(), a.mutable_data());
  return a;
}

py::list intList(const std::vector<int>& v) {
  py::list out;
  for (int x : v) out.append(x);
  return out;
}

void readSpot(const py::object& board, const py::object& oop, const py::object& ip, int64_t pot,
              int64_t stack, const py::object& bets, const py::object& raises, int maxRaises,
              pps::gpu::PostflopSpot& spot, pps::pf::BetConfig& cfg) {
  const std::vector<Card> bd = readBoard(board);
  spot.nBoard = (int)bd.size();
  for (int i = 0; i < spot.nBoard; ++i) spot.board[i] = bd[i];
  spot.oop = readRange(oop);
  spot.ip = readRange(ip);
  spot.pot = pot;
  spot.stack = stack;
  spot.oop.filterBoard(spot.board, spot.nBoard);
  spot.ip.filterBoard(spot.board, spot.nBoard);
  cfg.betFracs.clear();
  readBets(bets, cfg);
  cfg.raiseMults = readDoubles(raises);
  cfg.maxRaises = maxRaises;
  if (cfg.betFracs.empty() && !cfg.betAllIn && !cfg.betGeometric)
    throw std::runtime_error("empty bet abstraction: pass at least one size (e.g. \"0.75,a\")");
}

py::dict jobDict(const pps::gpu::JobInfo& j) {
  py::dict d;
  d["running"] = j.running;
  d["status"] = j.status.empty() ? py::object(py::none()) : py::object(py::str(j.status));
  d["kind"] = j.kind.empty() ? py::object(py::none()) : py::object(py::str(j.kind));
  d["start"] = j.start;
  d["done"] = j.done;
  d["total"] = j.total;
  d["target"] = j.target > 0.0 ? py::object(py::float_(j.target)) : py::object(py::none());
  d["error"] = j.error.empty() ? py::object(py::none()) : py::object(py::str(j.error));
  return d;
}

// A solver with its own engine, or a spot served by a SpotServer (then every
// call runs between the server's batched steps, and solves are jobs of the
// server, batched with the other spots' jobs).
class Solver {
 public:
  Solver(const py::object& board, const py::object& oop, const py::object& ip, int64_t pot,
         int64_t stack, const py::object& bets, const py::object& raises, int maxRaises,
         bool isomorphism) {
    pps::gpu::PostflopSpot spot;
    pps::pf::BetConfig cfg;
    readSpot(board, oop, ip, pot, stack, bets, raises, maxRaises, spot, cfg);
    pps::gpu::SolverOptions opt;
    opt.suitIsomorphism = isomorphism;
    py::gil_scoped_release nogil;
    s_ = std::make_unique<GpuPostflopSolver>(spot, cfg, opt);
  }

  Solver(std::shared_ptr<SpotServer> srv, int id) : srv_(std::move(srv)), id_(id) {}

  ~Solver() {
    if (!srv_) return;
    try {
      py::gil_scoped_release nogil;
      srv_->remove(id_);
    } catch (...) {
    }
  }

  static std::unique_ptr<Solver> load(const std::string& path) {
    std::unique_ptr<Solver> s(new Solver());
    py::gil_scoped_release nogil;
    s->s_ = std::make_unique<GpuPostflopSolver>(path);
    return s;
  }

  // Runs f(solver) without the GIL: directly, or on the server between steps.
  template <class F>
  auto on(F&& f) const -> decltype(f(std::declval<GpuPostflopSolver&>())) {
    using R = decltype(f(std::declval<GpuPostflopSolver&>()));
    py::gil_scoped_release nogil;
    if (!srv_) return f(*s_);
    if constexpr (std::is_void_v<R>) {
      srv_->with(id_, [&](GpuPostflopSolver& g) { f(g); });
    } else {
      std::optional<R> r;
      srv_->with(id_, [&](GpuPostflopSolver& g) { r.emplace(f(g)); });
      return std::move(*r);
    }
  }

  void solve(int maxIters, const std::string& algo, double target) {
    if (!srv_) return on([&](GpuPostflopSolver& g) { g.solve(maxIters, algo, target); });
    startSolve(maxIters, algo, target);
    finishJob();
  }
  void continueSolve(int maxIters, double target) {
    if (!srv_) return on([&](GpuPostflopSolver& g) { g.continueSolve(maxIters, target); });
    startContinue(maxIters, target);
    finishJob();
  }
  void reset() {
    if (!srv_) return on([](GpuPostflopSolver& g) { g.reset(); });
    py::gil_scoped_release nogil;
    srv_->reset(id_);
  }

  // ---- jobs of a served spot (no reference counterpart in the module; the
  // reference's HTTP layer runs them on threads, python/pps_api.py:350-403)
  py::dict startSolve(int maxIters, const std::string& algo, double target) {
    served("start_solve");
    {
      py::gil_scoped_release nogil;
      srv_->solve(id_, maxIters, algo, target);
    }
    return job();
  }
  py::dict startContinue(int maxIters, double target) {
    served("start_continue");
    {
      py::gil_scoped_release nogil;
      srv_->continueSolve(id_, maxIters, target);
    }
    return job();
  }
  py::dict job() {
    served("job");
    pps::gpu::JobInfo j;
    {
      py::gil_scoped_release nogil;
      j = srv_->job(id_);
    }
    return jobDict(j);
  }
  bool cancel() {
    served("cancel");
    py::gil_scoped_release nogil;
    return srv_->cancel(id_);
  }
  py::dict wait() {
    served("wait");
    pps::gpu::JobInfo j;
    {
      py::gil_scoped_release nogil;
      j = srv_->wait(id_);
    }
    return jobDict(j);
  }
  int id() const { return srv_ ? id_ : -1; }

  py::dict stats() {
    const pps::pf::NodeStats st = on([](GpuPostflopSolver& g) { return g.stats(); });
    py::dict d;
    d["ev_oop"] = st.ev0;
    d["ev_ip"] = st.ev1;
    d["exploitability"] = st.expl;
    d["pair_mass"] = st.pairMass;
    return d;
  }

  int numNodes() const { return on([](GpuPostflopSolver& g) { return g.numNodes(); }); }
  int maxDepth() const { return on([](GpuPostflopSolver& g) { return g.maxDepth(); }); }
  int numDecideNodes() const { return on([](GpuPostflopSolver& g) { return g.numDecideNodes(); }); }
  int iterationsRun() const { return on([](GpuPostflopSolver& g) { return g.iterationsRun(); }); }
  int64_t totalIterations() const {
    return on([](GpuPostflopSolver& g) { return g.totalIterations(); });
  }
  double iterationBytes() const { return on([](GpuPostflopSolver& g) { return g.iterationBytes(); }); }
  double lastSolveDeviceMs() const {
    return on([](GpuPostflopSolver& g) { return g.lastSolveDeviceMs(); });
  }

  py::list decideNodes() const {
    std::vector<pps::gpu::DecideNodeInfo> ns;
    on([&](GpuPostflopSolver& g) {
      for (int i = 0; i < g.numDecideNodes(); ++i) ns.push_back(g.decideNode(i));
    });
    py::list out;
    for (size_t i = 0; i < ns.size(); ++i) {
      py::dict d;
      d["index"] = i;
      d["node_id"] = ns[i].nodeId;
      d["player"] = ns[i].player;
      d["depth"] = ns[i].depth;
      d["n_board"] = ns[i].nBoard;
      d["n_combos"] = ns[i].nCombos;
      py::list acts;
      for (const auto& a : ns[i].actions) acts.append(actionDict(a));
      d["actions"] = acts;
      d["children"] = intList(ns[i].children);
      out.append(d);
    }
    return out;
  }

  py::dict strategy(int idx) {
    pps::gpu::ComboStrategy cs;
    pps::gpu::DecideNodeInfo n;
    std::vector<uint8_t> cards;
    std::vector<double> agg;
    on([&](GpuPostflopSolver& g) {
      cs = g.comboStrategy(idx);
      n = g.decideNode(idx);
      cards = g.playerCards(n.player);
      agg = g.nodeStrategy(idx);
    });
    const py::ssize_t na = (py::ssize_t)cs.actions.size(), nc = (py::ssize_t)cs.combos.size();
    py::dict d;
    d["index"] = idx;
    d["node_id"] = n.nodeId;
    d["player"] = n.player;
    d["depth"] = n.depth;
    py::list acts, labels;
    for (const auto& a : cs.actions) acts.append(actionDict(a));
    d["actions"] = acts;
    for (int c : cs.combos) labels.append(cardText(cards[2 * c]) + cardText(cards[2 * c + 1]));
    d["cards"] = labels;
    py::array_t<double> f({na, nc});
    std::copy(cs.freqs.begin(), cs.freqs.end(), f.mutable_data());
    d["freqs"] = f;
    d["combos"] = cs.combos;
    if (!agg.empty()) d["aggregate"] = arr(agg);
    return d;
  }

  py::dict rootEv() {
    pps::gpu::ComboEv ev;
    std::vector<uint8_t> cards[2];
    on([&](GpuPostflopSolver& g) {
      ev = g.rootEvPerCombo();
      for (int p = 0; p < 2; ++p) cards[p] = g.playerCards(p);
    });
    py::dict out;
    const char* names[2] = {"oop", "ip"};
    for (int p = 0; p < 2; ++p) {
      py::list labels;
      for (int c : ev.combos[p]) labels.append(cardText(cards[p][2 * c]) + cardText(cards[p][2 * c + 1]));
      py::dict d;
      d["combos"] = intList(ev.combos[p]);
      d["cards"] = labels;
      d["ev"] = arr(ev.ev[p]);
      d["mass"] = arr(ev.mass[p]);
      out[names[p]] = d;
    }
    return out;
  }

  py::dict nodeEv(int idx) {
    pps::gpu::NodeEv ev;
    std::vector<uint8_t> cards[2];
    on([&](GpuPostflopSolver& g) {
      ev = g.nodeEv(idx);
      for (int p = 0; p < 2; ++p) cards[p] = g.playerCards(p);
    });
    py::dict out;
    out["decider"] = ev.decider;
    py::list acts, aev;
    for (const auto& a : ev.actions) acts.append(actionDict(a));
    out["actions"] = acts;
    for (const auto& a : ev.actionEv) {
      py::dict d;
      d["per_combo"] = arr(a.perCombo);
      d["agg_ev"] = a.aggEv;
      aev.append(d);
    }
    out["action_ev"] = aev;
    const char* names[2] = {"oop", "ip"};
    py::dict sides;
    for (int p = 0; p < 2; ++p) {
      py::list labels;
      for (int c : ev.side[p].combos) labels.append(cardText(cards[p][2 * c]) + cardText(cards[p][2 * c + 1]));
      py::dict d;
      d["combos"] = intList(ev.side[p].combos);
      d["cards"] = labels;
      d["ev"] = arr(ev.side[p].ev);
      d["mass"] = arr(ev.side[p].mass);
      d["agg_ev"] = ev.side[p].aggEv;
      sides[names[p]] = d;
    }
    out["sides"] = sides;
    return out;
  }

  py::dict runoutSummary(int idx, int action) {
    const pps::gpu::RunoutSummary rs =
        on([&](GpuPostflopSolver& g) { return g.runoutSummary(idx, action); });
    py::dict out;
    out["decider"] = rs.decider;
    py::list acts, brs;
    for (const auto& a : rs.actions) acts.append(actionDict(a));
    out["actions"] = acts;
    for (const auto& b : rs.branches) {
      py::dict d, evs;
      d["card"] = b.card >= 0 ? cardText(b.card) : std::string("?");
      d["next_decide"] = b.nextDecide;
      d["freq"] = arr(b.freq);
      evs["oop"] = b.aggEv[0];
      evs["ip"] = b.aggEv[1];
      d["agg_ev"] = evs;
      brs.append(d);
    }
    out["branches"] = brs;
    return out;
  }

  void save(const std::string& path) const {
    on([&](GpuPostflopSolver& g) { g.save(path); });
  }

  py::list playerCards(int p) const {
    py::list out;
    for (uint8_t c : on([&](GpuPostflopSolver& g) { return g.playerCards(p); })) out.append((int)c);
    return out;
  }
  py::array_t<double> playerWeights(int p) const {
    return arr(on([&](GpuPostflopSolver& g) { return g.playerWeights(p); }));
  }
  int numBaseCombos(int p) const { return on([&](GpuPostflopSolver& g) { return g.numBaseCombos(p); }); }

  py::dict treeStructure() const {
    const pps::gpu::TreeStructure t = on([](GpuPostflopSolver& g) { return g.treeStructure(); });
    py::list kinds, base, kids, boards, nboard;
    for (uint8_t k : t.kinds) kinds.append((int)k);
    for (int c : t.childBase) base.append(c);
    for (int c : t.children) kids.append(c);
    for (uint8_t c : t.boards) boards.append((int)c);
    for (uint8_t c : t.nBoard) nboard.append((int)c);
    py::dict d;
    d["kinds"] = kinds;
    d["child_base"] = base;
    d["children"] = kids;
    d["boards"] = boards;
    d["n_board"] = nboard;
    return d;
  }

 private:
  Solver() = default;
  std::unique_ptr<GpuPostflopSolver> s_;
  std::shared_ptr<SpotServer> srv_;
  int id_ = -1;

  void served(const char* what) const {
    if (!srv_) throw std::runtime_error(std::string(what) + ": this solver is not served (use pps.Server.add)");
  }
  // blocks until the spot's job ends; a job error is raised here
  void finishJob() {
    const py::dict j = wait();
    if (!j["error"].is_none()) throw std::runtime_error(j["error"].cast<std::string>());
  }
};

// pps.Server: owns a SpotServer; hands out served Solvers.
class Server {
 public:
  Server(int64_t arenaBytes, int maxSpots, int maxStep) {
    pps::gpu::ServerOptions o;
    o.arenaBytes = arenaBytes;
    o.maxSpots = maxSpots;
    o.maxStep = maxStep;
    py::gil_scoped_release nogil;
    srv_ = std::make_shared<SpotServer>(o);
  }

  std::unique_ptr<Solver> add(const py::object& board, const py::object& oop, const py::object& ip,
                              int64_t pot, int64_t stack, const py::object& bets,
                              const py::object& raises, int maxRaises, bool isomorphism) {
    pps::gpu::PostflopSpot spot;
    pps::pf::BetConfig cfg;
    readSpot(board, oop, ip, pot, stack, bets, raises, maxRaises, spot, cfg);
    pps::gpu::SolverOptions opt;
    opt.suitIsomorphism = isomorphism;
    int id = -1;
    {
      py::gil_scoped_release nogil;
      id = srv_->add(spot, cfg, opt);
    }
    return std::make_unique<Solver>(srv_, id);
  }

  std::unique_ptr<Solver> load(const std::string& path) {
    int id = -1;
    {
      py::gil_scoped_release nogil;
      id = srv_->load(path);
    }
    return std::make_unique<Solver>(srv_, id);
  }

  py::dict stats() {
    pps::gpu::ServerStats s;
    {
      py::gil_scoped_release nogil;
      s = srv_->stats();
    }
    py::dict d;
    d["steps"] = s.steps;
    d["iterations"] = s.iterations;
    d["device_ms"] = s.deviceMs;
    d["check_ms"] = s.checkMs;
    d["resident"] = s.resident;
    d["running"] = s.running;
    d["peak_running"] = s.peakRunning;
    d["arena_bytes"] = s.arenaBytes;
    d["free_bytes"] = s.freeBytes;
    return d;
  }

 private:
  std::shared_ptr<SpotServer> srv_;
};

}  // namespace

PYBIND11_MODULE(pps_native, m) {
  m.doc() = "MI355X postflop solver (DCFR; compiled tree, HIP kernels, hipGraph replay)";
  py::class_<Solver>(m, "Solver")
      .def(py::init([](const py::object& board, const py::object& oop, const py::object& ip,
                       int64_t pot, int64_t stack, const py::object& bets,
                       const py::object& raises, int maxRaises, bool isomorphism) {
             return std::make_unique<Solver>(board, oop, ip, pot, stack, bets, raises, maxRaises,
                                             isomorphism);
           }),
           py::arg("board"), py::arg("oop"), py::arg("ip"), py::arg("pot"), py::arg("stack"),
           py::arg("bets") = "0.75,a", py::arg("raises") = "2.5,3", py::arg("max_raises") = 0,
           py::arg("isomorphism") = false)
      .def_static("load", &Solver::load, py::arg("path"))
      .def("solve", &Solver::solve, py::arg("max_iters"), py::arg("algo") = "dcfr",
           py::arg("target") = -1.0)
      .def("continue_solve", &Solver::continueSolve, py::arg("max_iters"), py::arg("target") = -1.0)
      .def("reset", &Solver::reset)
      .def("stats", &Solver::stats)
      .def_property_readonly("num_nodes", &Solver::numNodes)
      .def_property_readonly("max_depth", &Solver::maxDepth)
      .def_property_readonly("num_decide_nodes", &Solver::numDecideNodes)
      .def_property_readonly("iterations_run", &Solver::iterationsRun)
      .def_property_readonly("total_iterations", &Solver::totalIterations)
      .def_property_readonly("iteration_bytes", &Solver::iterationBytes)
      .def_property_readonly("last_solve_device_ms", &Solver::lastSolveDeviceMs)
      .def("decide_nodes", &Solver::decideNodes)
      .def("strategy", &Solver::strategy, py::arg("decide_idx") = 0)
      .def("root_ev", &Solver::rootEv)
      .def("node_ev", &Solver::nodeEv, py::arg("decide_idx") = 0)
      .def("runout_summary", &Solver::runoutSummary, py::arg("decide_idx"), py::arg("action_idx"))
      .def("save", &Solver::save, py::arg("path"))
      .def("player_cards", &Solver::playerCards, py::arg("player"))
      .def("player_weights", &Solver::playerWeights, py::arg("player"))
      .def("num_base_combos", &Solver::numBaseCombos, py::arg("player"))
      .def("tree_structure", &Solver::treeStructure)
      .def("start_solve", &Solver::startSolve, py::arg("max_iters"), py::arg("algo") = "dcfr",
           py::arg("target") = -1.0)
      .def("start_continue", &Solver::startContinue, py::arg("max_iters"), py::arg("target") = -1.0)
      .def("job", &Solver::job)
      .def("cancel", &Solver::cancel)
      .def("wait", &Solver::wait)
      .def_property_readonly("spot_id", &Solver::id);
  py::class_<Server>(m, "Server")
      .def(py::init<int64_t, int, int>(), py::arg("arena_bytes") = 0, py::arg("max_spots") = 256,
           py::arg("max_step") = 32)
      .def("add", &Server::add, py::arg("board"), py::arg("oop"), py::arg("ip"), py::arg("pot"),
           py::arg("stack"), py::arg("bets") = "0.75,a", py::arg("raises") = "2.5,3",
           py::arg("max_raises") = 0, py::arg("isomorphism") = false)
      .def("load", &Server::load, py::arg("path"))
      .def("stats", &Server::stats);
  py::register_exception<pps::gpu::ArenaFull>(m, "ArenaFull", PyExc_MemoryError);
}

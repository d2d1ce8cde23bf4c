// Solution files: one schema (describe) walked in both directions, by an
// encoder that writes each field and by a decoder that reads, bounds-checks
// and validates it. The field order and encodings are the PPS1 v1 layout of
// the reference (cuda/gpu_cfr.cu:2686-2853); the checks are this build's.
#include "solution_file.h"

#include <sys/types.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

namespace pps {
namespace gpu {

namespace {

constexpr uint32_t kPps1Magic = 0x31535050u;  // the bytes "PPS1"
constexpr uint32_t kPps1Version = 1;
constexpr int32_t kMaxSizes = 32;  // a decide node has at most 32 actions
constexpr int32_t kMaxAlgoName = 32;

bool knownAlgorithm(const std::string& a) { return a == "dcfr" || a == "dcfr-g2" || a == "hs30"; }

std::string num(double x) {
  char b[40];
  std::snprintf(b, sizeof(b), "%.17g", x);
  return b;
}

class Encoder {
 public:
  static constexpr bool kReading = false;
  Encoder(FILE* f, const std::string& path) : f_(f), path_(path) {}
  template <typename T>
  void scalar(const char*, T& v) {
    put(&v, sizeof(T));
  }
  void flag(const char*, bool& v) {
    const int8_t b = v ? 1 : 0;
    put(&b, 1);
  }
  void bytes(const char*, void* p, size_t n) { put(p, n); }
  void sizes(const char*, std::vector<double>& v) {
    const int32_t n = (int32_t)v.size();
    put(&n, sizeof(n));
    put(v.data(), v.size() * sizeof(double));
  }
  void name(const char*, std::string& s) {
    const int32_t n = (int32_t)s.size();
    put(&n, sizeof(n));
    put(s.data(), s.size());
  }
  void table(const char*, std::vector<float>& v, int64_t n) { put(v.data(), (size_t)n * sizeof(float)); }
  [[noreturn]] void fail(const char*, const std::string&) { throw std::logic_error("encoder check"); }

 private:
  void put(const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f_) != n)
      throw std::runtime_error("solution file '" + path_ + "': write failed");
  }
  FILE* f_;
  const std::string& path_;
};

class Decoder {
 public:
  static constexpr bool kReading = true;
  Decoder(FILE* f, const std::string& path, int64_t size, bool withRows)
      : f_(f), path_(path), size_(size), withRows_(withRows) {}
  template <typename T>
  void scalar(const char* field, T& v) {
    get(field, &v, sizeof(T));
  }
  void flag(const char* field, bool& v) {
    int8_t b = 0;
    get(field, &b, 1);
    if (b != 0 && b != 1) fail(field, "byte " + std::to_string((int)b) + " is not a flag (0 or 1)");
    v = b != 0;
  }
  void bytes(const char* field, void* p, size_t n) { get(field, p, n); }
  void sizes(const char* field, std::vector<double>& v) {
    int32_t n = 0;
    get(field, &n, sizeof(n));
    if (n < 0 || n > kMaxSizes)
      fail(field, std::to_string(n) + " entries (a list holds 0 to " + std::to_string(kMaxSizes) + ")");
    v.assign((size_t)n, 0.0);
    get(field, v.data(), v.size() * sizeof(double));
    for (double x : v)
      if (!(std::isfinite(x) && x > 0.0)) fail(field, "entry " + num(x) + " is not a positive size");
  }
  void name(const char* field, std::string& s) {
    int32_t n = 0;
    get(field, &n, sizeof(n));
    if (n <= 0 || n > kMaxAlgoName)
      fail(field, "name length " + std::to_string(n) + " (1 to " + std::to_string(kMaxAlgoName) + ")");
    s.assign((size_t)n, '\0');
    get(field, &s[0], (size_t)n);
  }
  void table(const char* field, std::vector<float>& v, int64_t n) {
    const int64_t bytes = n * (int64_t)sizeof(float);
    need(field, bytes);
    if (!withRows_) {
      if (fseeko(f_, (off_t)bytes, SEEK_CUR) != 0) fail(field, "seek failed");
      pos_ += bytes;
      return;
    }
    v.assign((size_t)n, 0.0f);
    get(field, v.data(), (size_t)bytes);
  }
  [[noreturn]] void fail(const char* field, const std::string& problem) const {
    throw std::runtime_error("solution file '" + path_ + "': " + field + " at byte " +
                             std::to_string(pos_) + ": " + problem);
  }
  void atEnd() const {
    if (pos_ != size_)
      fail("end of data", std::to_string(size_ - pos_) + " more bytes follow the strategy rows");
  }

 private:
  void need(const char* field, int64_t n) const {
    if (n < 0 || n > size_ - pos_)
      fail(field, "needs " + std::to_string(n) + " bytes, the file has " +
                      std::to_string(size_ - pos_) + " left");
  }
  void get(const char* field, void* p, size_t n) {
    need(field, (int64_t)n);
    if (n && std::fread(p, 1, n, f_) != n) fail(field, "read failed");
    pos_ += (int64_t)n;
  }
  FILE* f_;
  const std::string& path_;
  int64_t size_, pos_ = 0;
  bool withRows_;
};

// A check applied when reading (messages built only on failure).
template <typename IO, typename Msg>
void require(IO& io, bool ok, const char* field, Msg&& message) {
  if constexpr (IO::kReading)
    if (!ok) io.fail(field, message());
}

// The PPS1 v1 fields in file order, with the reader's checks.
template <typename IO>
void describe(IO& io, Solution& s) {
  uint32_t magic = kPps1Magic, version = kPps1Version;
  io.scalar("magic", magic);
  require(io, magic == kPps1Magic, "magic", [] { return std::string("not a PPS1 solution file"); });
  io.scalar("version", version);
  require(io, version == kPps1Version, "version", [&] {
    return "version " + std::to_string(version) + " (this build reads version 1)";
  });
  int32_t nBoard = s.spot.nBoard;
  io.scalar("board size", nBoard);
  require(io, nBoard >= 3 && nBoard <= 5, "board size",
          [&] { return std::to_string(nBoard) + " cards (a flop, turn or river board has 3 to 5)"; });
  s.spot.nBoard = nBoard;
  io.bytes("board", s.spot.board, sizeof(s.spot.board));
  for (int i = 0; i < nBoard; ++i) {
    require(io, s.spot.board[i] < 52, "board",
            [&] { return "card id " + std::to_string((int)s.spot.board[i]) + " (0 to 51)"; });
    for (int j = 0; j < i; ++j)
      require(io, s.spot.board[i] != s.spot.board[j], "board",
              [&] { return "card " + std::to_string((int)s.spot.board[i]) + " appears twice"; });
  }
  io.scalar("pot", s.spot.pot);
  require(io, s.spot.pot > 0, "pot", [&] { return std::to_string(s.spot.pot) + " chips"; });
  io.scalar("stack", s.spot.stack);
  require(io, s.spot.stack >= 0, "stack", [&] { return std::to_string(s.spot.stack) + " chips"; });
  for (int p = 0; p < 2; ++p) {
    const char* field = p == 0 ? "OOP range" : "IP range";
    double* w = (p == 0 ? s.spot.oop : s.spot.ip).w;
    io.bytes(field, w, sizeof(double) * kCombos);
    for (int k = 0; k < kCombos; ++k)
      require(io, std::isfinite(w[k]) && w[k] >= 0.0, field,
              [&] { return "combo " + std::to_string(k) + " has weight " + num(w[k]); });
  }
  io.sizes("bet sizes", s.cfg.betFracs);
  io.sizes("raise sizes", s.cfg.raiseMults);
  io.flag("all-in bet flag", s.cfg.betAllIn);
  io.flag("geometric bet flag", s.cfg.betGeometric);
  int32_t cap = s.cfg.maxRaises;
  io.scalar("raise cap", cap);
  require(io, cap >= 0, "raise cap", [&] { return std::to_string(cap) + " (0 = no cap)"; });
  s.cfg.maxRaises = cap;
  io.scalar("add-all-in threshold", s.cfg.addAllinThreshold);
  io.scalar("force-all-in threshold", s.cfg.forceAllinThreshold);
  io.scalar("merging threshold", s.cfg.mergingThreshold);
  require(io,
          std::isfinite(s.cfg.addAllinThreshold) && std::isfinite(s.cfg.forceAllinThreshold) &&
              std::isfinite(s.cfg.mergingThreshold),
          "thresholds", [] { return std::string("not finite"); });
  io.scalar("node count", s.nodes);
  require(io, s.nodes > 0, "node count", [&] { return std::to_string(s.nodes); });
  io.scalar("iterations", s.iterations);
  require(io, s.iterations >= 0 && s.iterations <= INT_MAX, "iterations",
          [&] { return std::to_string(s.iterations) + " (0 to 2^31 - 1)"; });
  io.name("algorithm", s.algo);
  require(io, knownAlgorithm(s.algo), "algorithm",
          [&] { return "'" + s.algo + "' (dcfr, dcfr-g2 or hs30)"; });
  io.scalar("row count", s.rows);
  require(io, s.rows >= 0, "row count", [&] { return std::to_string(s.rows); });
  io.table("regret rows", s.regret, s.rows);
  io.table("strategy rows", s.strat, s.rows);
}

}  // namespace

void writeSolution(const std::string& path, const Solution& s) {
  if ((int64_t)s.regret.size() != s.rows || (int64_t)s.strat.size() != s.rows)
    throw std::runtime_error("solution: row tables do not hold " + std::to_string(s.rows) + " entries");
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("solution file '" + path + "': cannot open for writing");
  try {
    Encoder enc(f, path);
    describe(enc, const_cast<Solution&>(s));  // the encoder only reads the fields
    if (std::fflush(f) != 0) throw std::runtime_error("solution file '" + path + "': write failed");
  } catch (...) {
    std::fclose(f);
    throw;
  }
  if (std::fclose(f) != 0) throw std::runtime_error("solution file '" + path + "': write failed");
}

Solution readSolution(const std::string& path, bool withRows) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("solution file '" + path + "': cannot open");
  Solution s;
  try {
    if (fseeko(f, 0, SEEK_END) != 0) throw std::runtime_error("solution file '" + path + "': seek failed");
    const int64_t size = (int64_t)ftello(f);
    if (size < 0 || fseeko(f, 0, SEEK_SET) != 0)
      throw std::runtime_error("solution file '" + path + "': seek failed");
    Decoder dec(f, path, size, withRows);
    describe(dec, s);
    dec.atEnd();
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  return s;
}

}  // namespace gpu
}  // namespace pps

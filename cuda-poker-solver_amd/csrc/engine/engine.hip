// MI355X DCFR engine: gfx950 kernels + the ppsg C ABI (include/ppsg.h).
//
// Memory: one device arena per engine. Every resident spot (compiled tree)
// owns one chunk of it (index tables + f32 rows), so one engine runs one
// spot (ppsg_create) or many (ppsg_batch_*) with the same kernels. Work
// items carry byte offsets into the arena.
//
// One DCFR iteration = two alternating half-steps (traverser 0, then 1);
// each half-step is a forward pass over depths 0..D (opponent reach, plus
// the fold/showdown values of each depth) and a backward pass over depths
// D..0 (traverser counterfactual values with the fused DCFR update). For a
// batch, depth d of every active spot goes into the same launch (a depth
// wavefront over uneven trees), so the launch count per iteration is set by
// the deepest tree, not by the number of spots. The whole iteration (plus
// the per-spot schedule-counter bumps) is captured once in a hipGraph and
// replayed; it is captured again only when the active set changes.
//
// Work items are 64-byte descriptors built on the host with every offset
// resolved: a workgroup issues its data loads after one (scalar) descriptor
// load. Sibling rows are contiguous (child k at base + k * stride).
//
//   forward  decide (opponent acts)  one lane per opponent combo, 256 per item
//   forward  chance (branch, tile)   child reach = parent reach / (52-board-4)
//   forward  fold / showdown         one workgroup per node: opponent reach
//                                    gathered + one blocked f64 LDS scan
//   backward decide (tile)           RM+ value + fused DCFR update, or the
//                                    opponent-side sum of child values
//   backward chance (tile)           sum of child values over surviving branches
// Traverser-acting decide nodes have no forward item: their children's
// reach rows alias the node's row (cuda/gpu_cfr.cu:1485-1514).
//
// No float atomics anywhere: every sum has a fixed order, so results are
// bitwise reproducible and do not depend on launch shape or on which other
// spots share the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ppsg.h"

namespace {

constexpr int kBlock = 256;  // 4 wave64s
constexpr int kWideBlock = 512;  // wide showdown groups: 8 wave64s
// Decide and chance work items are per wave: a 256-lane block runs four
// consecutive items, one per wave, so a node's last tile spans at most 63
// lanes of row padding (not 255) and the waves of a block may serve
// different nodes.
constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxList = 1326;
constexpr int kRowRegs = (kMaxList + kBlock - 1) / kBlock;  // 6
constexpr int64_t kAlign = 256;
constexpr int kLists = 5;  // per traverser: forward, backward, fold, showdown, wide showdown
// bytes: 53 int32 segment bounds, then the 52 cards in the order their
// lane groups take them (longest segment first), then pad
constexpr int kGatherHeader = 320;
constexpr int kCardOrderAt = 53 * 4;  // byte offset of the card order
// fold: entries of a card's segment per lane (4 lanes per card)
constexpr int kFoldPart = 12;
// a lane's part is gathered in chunks of this many entries; a wave skips the
// chunks past its longest part
constexpr int kPartChunk = 3;
constexpr int kMaxGroup = 64;  // terminal nodes per work item at most
constexpr int kGraphIters = 8;  // iterations per replay of a stable solve, at most
constexpr int kGraphNodes = 256;  // kernels per multi-iteration graph, at most
// Combos per lane (kWave apart) of forward decide, forward chance and
// backward decide items: a per-depth launch shape. Spots are planned at one
// combo per lane; a list built for a depth merges each node's tiles to the
// depth's shape. Defaults (measured over the whole batched stream: more
// loads in flight per lane help the read-mostly forward items, the
// read-modify-write backward tiles do best with one combo per lane);
// ppsg_autotune_shapes picks per depth from the candidates below.
constexpr int kRDecide = 2;
constexpr int kRBackward = 1;
constexpr int kRChance = 4;
constexpr int kRDecideCand[] = {1, 2, 4};
constexpr int kRChanceCand[] = {2, 4, 8};
constexpr int kRBackwardCand[] = {1, 2};

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return -1;
}

#define HIP_TRY(call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(std::string(#call) + ": " + hipGetErrorString(e_) + " (" __FILE__ ":" + \
                  std::to_string(__LINE__) + ")");                                        \
  } while (0)

// Per-spot control block: the device-side schedule index and table, the
// running iteration's coefficients, and whether the spot's work items still
// run (retired spots stay in the merged work lists until the next
// compaction, masked out by `alive`).
struct alignas(32) Ctl {
  long long counter;  // schedule index of the running iteration
  const float* coef;  // 3 floats per schedule index (+ one zero entry)
  float cur[3];       // coef[3 * counter ..]: positive, negative, average
  int alive;
};
static_assert(sizeof(Ctl) == 32, "control block is one 32-byte load");

enum ItemKind : uint8_t {
  kFwdDecide = 0,
  kFwdChance = 1,
  kBwdDecide = 2,
  kBwdChance = 3,
  kBwdFold = 4,
  kBwdShowdown = 5,
};

// One work item; a, b, c, d are byte offsets into the arena; f is the
// owning spot's slot (its control block).
//   kFwdDecide   a regret row, d strategy row (action rows ld apart), b node
//                reach row, c children's reach block, m child stride
//                (floats), n opponent combos, na actions
//   kFwdChance   a to-parent map (u16), b parent reach row, c child reach
//                row, n child opponent combos, v0 = 1/(52-board-4)
//   (decide and chance items are one wave's work: combos [tile * kWave * r,
//   +kWave * r), lane t taking t, t + kWave, ... so that r loads per row
//   are in flight)
//   kBwdDecide   a regret row, d strategy row (action rows ld apart), b node
//                cfv row, c children's cfv block, m child stride (floats), n
//                traverser combos, e = actor is the traverser
//   kBwdChance   a dealt-card x parent-slot map (i16: child slot or -1), b
//                node cfv row, c children's cfv block, d the child serving
//                each dealt card (u8), m child stride, n traverser combos
//                (parent), na dealt cards, e = every card is served by its
//                own child (d is not read)
//   kBwdFold /   a group of fold (or showdown) nodes of one board class:
//   kBwdShowdown a the group's first TermNode, cnt nodes in the group, b the
//                spot's chunk (TermNode offsets are relative to it), c the
//                opponent's gather block of the class (53 i32 per-card
//                segment bounds, then the u16 gather list: showdown, slots
//                in strength order + the per-card segments; fold, the
//                per-card segments), d per traverser combo metadata (fold: u32 cards
//                + same; showdown: 4 u32), na row registers per lane (3 or
//                6), n traverser combos, m opponent combos
struct alignas(16) Item {
  int64_t a, b, c, d;
  int32_t n, m, e, f;
  uint8_t kind, na;
  uint8_t tile;  // decide / chance: which run of kWave * r combos
  uint8_t r;     // decide / chance: combos per lane (kWave apart)
  float v0;
  int32_t cnt;   // terminal groups: nodes in the group
  int32_t ld;    // decide: distance between action rows (floats)
};
static_assert(sizeof(Item) == 64, "work item must be one 64-byte line");

// One fold or showdown node of a terminal group: its opponent reach row and
// its cfv row (byte offsets in the spot's chunk) and payoffs (fold: v0 per
// unit of opponent mass; showdown: win, tie, lose).
struct alignas(16) TermNode {
  int64_t reach, out;
  float v0, v1, v2, pad;
};
static_assert(sizeof(TermNode) == 32, "terminal node entry is 32 bytes");

struct Dev {
  char* arena;
  Ctl* ctl;
};

template <typename T>
__device__ __forceinline__ T* at(const Dev& d, int64_t off) {
  return reinterpret_cast<T*>(d.arena + off);
}

// Per-wave work items: this wave's item index (a scalar, so the item's
// fields are scalar loads) and the lane within the wave.
__device__ __forceinline__ int waveItem() {
  return (int)blockIdx.x * kWavesPerBlock + __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
}
__device__ __forceinline__ int waveLane() { return (int)threadIdx.x & (kWave - 1); }

typedef int v4i __attribute__((ext_vector_type(4)));

// Item w (clamped into the list, so the load needs no branch first), all
// 64 bytes in one round trip: the empty asm takes every word, so none of
// the loads is split off behind the kind test; the arena bases are taken
// with the list arguments.
__device__ __forceinline__ Item fetchItem(const Dev& d, const Item* __restrict__ items, int w,
                                          int count) {
  asm volatile("" : : "s"(d.arena), "s"(d.ctl));
  const v4i* p = reinterpret_cast<const v4i*>(items + min(w, count - 1));
  v4i q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
  asm volatile("" : "+s"(q0), "+s"(q1), "+s"(q2), "+s"(q3));
  Item it;
  v4i* o = reinterpret_cast<v4i*>(&it);
  o[0] = q0;
  o[1] = q1;
  o[2] = q2;
  o[3] = q3;
  return it;
}

__device__ __forceinline__ float sigmaOf(float x, float acc, float uniform, bool plus) {
  if (plus) return acc > 0.0f ? (x > 0.0f ? x / acc : 0.0f) : uniform;
  return acc > 0.0f ? x / acc : uniform;
}

// ---------------------------------------------------------------------------
// Forward reach.
// ---------------------------------------------------------------------------

// The end of the whole 128-byte lines (32 floats) that hold a row's first n
// values. Rows are padded to PPSG_ROW_STRIDE(n) floats (whole 64-lane
// tiles); a tile's lanes up to lineEnd(n) load (indices clamped into the
// row's last needed line, so no load sits behind a per-lane branch) and
// store, the lanes past it store nothing. Every store then covers whole
// cache lines: a store that covers part of a line costs far more than its
// share (a launch whose nodes ended in part-filled lines ran ~15 % slower),
// while the lanes past the last line only moved padding (~1 % of a
// backward launch's time).
__host__ __device__ constexpr int lineEnd(int n) { return (n + 31) & ~31; }

// Decide (opponent acts): child reach a = reach * sigma_a. NA > 0: action
// count fixed at compile time (the row values stay in registers); every
// lane handles up to R combos kWave apart, all loads issued first.
template <int NA, int R>
__device__ __forceinline__ void fwdDecide(const Dev& d, const Item& it, int fromStrat) {
  const int c0 = it.tile * kWave * R + waveLane();
  const int n = lineEnd(it.n), na = NA > 0 ? NA : it.na;
  if (c0 >= n) return;
  const int64_t ld = it.ld, cs = it.m;
  const float* __restrict__ reach = at<float>(d, it.b);
  const float* __restrict__ src = at<float>(d, fromStrat ? it.d : it.a);
  float* __restrict__ out = at<float>(d, it.c);
  const bool plus = !fromStrat;
  const float uniform = 1.0f / (float)na;
  if (NA > 0) {
    float x[R][NA > 0 ? NA : 1], r[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int c = min(c0 + k * kWave, n - 1);
      r[k] = reach[c];
#pragma unroll
      for (int a = 0; a < NA; ++a) x[k][a] = src[a * ld + c];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int c = c0 + k * kWave;
      if (c >= n) continue;
      float acc = 0.0f;
#pragma unroll
      for (int a = 0; a < NA; ++a) acc += plus ? (x[k][a] > 0.0f ? x[k][a] : 0.0f) : x[k][a];
#pragma unroll
      for (int a = 0; a < NA; ++a) out[a * cs + c] = r[k] * sigmaOf(x[k][a], acc, uniform, plus);
    }
    return;
  }
  for (int k = 0; k < R; ++k) {
    const int c = c0 + k * kWave;
    if (c >= n) return;
    const float rr = reach[c];
    float acc = 0.0f;
    for (int a = 0; a < na; ++a) {
      const float x = src[a * ld + c];
      acc += plus ? (x > 0.0f ? x : 0.0f) : x;
    }
    for (int a = 0; a < na; ++a) out[a * cs + c] = rr * sigmaOf(src[a * ld + c], acc, uniform, plus);
  }
}

template <int RD, int RC>
__device__ __forceinline__ void fwdItem(const Dev& d, const Item& it, int fromStrat) {
  if (it.kind == kFwdChance) {
    const int c0 = it.tile * kWave * RC + waveLane();
    const float* __restrict__ in = at<float>(d, it.b);
    const uint16_t* __restrict__ tp = at<uint16_t>(d, it.a);
    float* __restrict__ out = at<float>(d, it.c);
    // whole lines: zeros into the row's last line past n; the loads are
    // issued for every lane (index clamped into the list) and the lanes
    // past n drop the value, so no load sits behind a per-lane branch
    const int n = it.n, nw = lineEnd(n);
    float x[RC];
#pragma unroll
    for (int k = 0; k < RC; ++k) x[k] = in[tp[min(c0 + k * kWave, n - 1)]];
#pragma unroll
    for (int k = 0; k < RC; ++k)
      if (c0 + k * kWave < nw) out[c0 + k * kWave] = c0 + k * kWave < n ? x[k] * it.v0 : 0.0f;
    return;
  }
  switch (it.na) {
    case 1: fwdDecide<1, RD>(d, it, fromStrat); break;
    case 2: fwdDecide<2, RD>(d, it, fromStrat); break;
    case 3: fwdDecide<3, RD>(d, it, fromStrat); break;
    case 4: fwdDecide<4, RD>(d, it, fromStrat); break;
    case 5: fwdDecide<5, RD>(d, it, fromStrat); break;
    case 6: fwdDecide<6, RD>(d, it, fromStrat); break;
    default: fwdDecide<0, RD>(d, it, fromStrat); break;
  }
}

// ---------------------------------------------------------------------------
// Fold and showdown terminals, one workgroup per group of nodes that share a
// board class (and so the class's index tables, which stay in registers
// across the group).
// ---------------------------------------------------------------------------

// Loads node g's opponent reach row (ROWS floats per lane, NT lanes) when
// g < cnt; zeros when the class has no opponent combos (nO == 0).
template <int ROWS, int NT = kBlock>
__device__ __forceinline__ void loadReach(float (&rv)[ROWS], const char* base,
                                          const TermNode* nodes, int g, int cnt, int nO) {
  if (g >= cnt) return;
  if (nO <= 0) {  // a dealt card blocks the whole opposing list: no mass
#pragma unroll
    for (int q = 0; q < ROWS; ++q) rv[q] = 0.0f;
    return;
  }
  const float* __restrict__ r = reinterpret_cast<const float*>(base + nodes[g].reach);
  // unconditional loads (index clamped into the row; callers stage only
  // j < nO): no register is written while a load into it is in flight
#pragma unroll
  for (int q = 0; q < ROWS; ++q) rv[q] = r[min((int)threadIdx.x + q * NT, nO - 1)];
}

// The largest x over the wave, as a wave-uniform (scalar) value.
__device__ __forceinline__ int waveMax(int x) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) x = max(x, __shfl_xor(x, off, 64));
  return __builtin_amdgcn_readfirstlane(x);
}

// Gathers raw[idx[k]] for the chunks of a lane's card-segment part that the
// wave's longest part (partMax, uniform) reaches. The entries past a lane's
// own part hold the zero sentinel slot, so the skipped chunks are terms of
// +0.0 and every sum over sv is the same with or without them.
__device__ __forceinline__ void gatherPart(float (&sv)[kFoldPart], const float* raw,
                                           const uint16_t (&idx)[kFoldPart], int partMax) {
#pragma unroll
  for (int k = 0; k < kFoldPart; ++k) {
    if (k % kPartChunk == 0 && k >= partMax) break;
    sv[k] = raw[idx[k]];
  }
}

// A group's node descriptors, copied into LDS once: reading them from the
// arena inside the node loop would be vector loads ordered behind every
// output store in flight (the arena is also written), so the next node's
// reach request would wait for all of them.
__device__ __forceinline__ const TermNode* stageNodes(TermNode* grp, const TermNode* nodes,
                                                      int cnt) {
  const int tid = threadIdx.x;
  if (tid < 2 * cnt) {
    const int4* src = reinterpret_cast<const int4*>(nodes);
    reinterpret_cast<int4*>(grp)[tid] = src[tid];
  }
  __syncthreads();
  return grp;
}

// A node's results, stored once the next node's reach is staged: its
// stores are then older than the next reach request, so the wait for that
// reach (which waits for all vector memory operations) never waits for
// them. Stores run to padT = lineEnd(nT): waves past it skip with a scalar
// test, and the lanes between nT and padT store values that nothing reads,
// so every store covers whole cache lines.
template <int ROWS, int NT>
__device__ __forceinline__ void storePadded(float* out, const float (&res)[ROWS], int waveBase,
                                            int padT) {
#pragma unroll
  for (int q = 0; q < ROWS; ++q)
    if (waveBase + q * NT < padT && (int)threadIdx.x + q * NT < padT)
      out[(int)threadIdx.x + q * NT] = res[q];
}

// Fold group: mass(i) = total - cardSum[c1] - cardSum[c2] + r[same(i)].
// The class's per-card segments (opponent slots holding card c) are split
// over 4 lanes per card (lanes 4c..4c+3, up to kFoldPart slots each, padded
// with the zero sentinel slot); each lane sums its part in f64, the 4 parts
// combine pairwise, and the total is half the sum of the card sums. The
// next node's reach row is requested as soon as this one is staged.
template <int ROWS>
__device__ void foldGroup(const Dev& d, const Item& it, char* lds) {
  __shared__ double cs[53];
  __shared__ TermNode grp[kMaxGroup];
  const int tid = threadIdx.x, lane = tid & 63;
  const int waveBase = __builtin_amdgcn_readfirstlane(tid >> 6) * 64;  // uniform
  const int nO = it.m, nT = it.n, cnt = it.cnt;
  const int padT = lineEnd(nT);
  float* raw = reinterpret_cast<float*>(lds);
  const TermNode* nodes = stageNodes(grp, at<TermNode>(d, it.a), cnt);
  const char* base = d.arena + it.b;
  // class data for the whole group; every load is issued by every lane
  // (indices clamped into the tables) and the lanes that need no value drop
  // it: a load under a per-lane condition is compiled to a branch with its
  // own wait, which turned these 15 loads into 15 round trips per group
  uint32_t tc[ROWS];
  {
    const uint32_t* cards = at<uint32_t>(d, it.d);
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      const int i = tid + q * kBlock;
      const uint32_t v = cards[min(i, max(nT - 1, 0))];
      tc[q] = i < nT ? v : 0u;
    }
  }
  // this lane's part of its card's segment
  uint16_t idx[kFoldPart];
  int partMax;  // the wave's longest part
  // the card of this lane group (the groups take the cards longest segment
  // first); lane groups 52.. load card order[51] and drop it
  const int card = at<uint8_t>(d, it.c + kCardOrderAt)[min(tid >> 2, 51)];
  {
    const int c = card, part = tid & 3;
    const int32_t* so = at<int32_t>(d, it.c);
    const int s0 = so[c], sl = so[c + 1] - s0;
    const int lo = s0 + (sl * part) / 4;
    const int len = (tid >> 2) < 52 ? s0 + (sl * (part + 1)) / 4 - lo : 0;
    const uint16_t* list = at<uint16_t>(d, it.c + kGatherHeader);
    const int last = max(2 * nO - 1, 0);  // the per-card lists hold 2 nO slots
#pragma unroll
    for (int k = 0; k < kFoldPart; ++k) {
      const uint16_t e = list[min(lo + k, last)];
      idx[k] = k < len ? e : (uint16_t)nO;
    }
    partMax = waveMax(len);
  }
  float res[ROWS];
  float* pend = nullptr;  // the previous node's output row
  auto node = [&](int g, float(&rv)[ROWS]) {
    const TermNode nd = nodes[g];
    if (g > 0) __syncthreads();  // the previous node's queries are done
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      const int j = tid + q * kBlock;
      if (j < nO) raw[j] = rv[q];
    }
    if (tid == 0) raw[nO] = 0.0f;
    if (g > 0) storePadded<ROWS, kBlock>(pend, res, waveBase, padT);
    __syncthreads();
    loadReach<ROWS>(rv, base, nodes, g + 1, cnt, nO);  // the next node's reach
    float sv[kFoldPart];
    gatherPart(sv, raw, idx, partMax);
    double part = 0.0;
#pragma unroll
    for (int k = 0; k < kFoldPart; ++k) {
      if (k % kPartChunk == 0 && k >= partMax) break;
      part += (double)sv[k];
    }
    part += __shfl_xor(part, 1, 64);
    part += __shfl_xor(part, 2, 64);
    if ((tid & 3) == 0 && (tid >> 2) < 52) cs[card] = part;
    __syncthreads();
    if (tid < 64) {  // total = half the sum of the card sums
      double x = tid < 52 ? cs[tid] : 0.0;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
      if (lane == 0) cs[52] = 0.5 * x;
    }
    __syncthreads();
    pend = reinterpret_cast<float*>(d.arena + it.b + nd.out);
    const double tot = cs[52];
    const double v = nd.v0;
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      const int c1 = tc[q] & 0xff, c2 = (tc[q] >> 8) & 0xff;
      const int same = (int)(tc[q] >> 16) - 1;
      double m = tot - cs[c1] - cs[c2];
      if (same >= 0) m += (double)raw[same];
      if (m < 0.0) m = 0.0;
      res[q] = (float)(m * v);
    }
  };
  float rv[ROWS];
  loadReach<ROWS>(rv, base, nodes, 0, cnt, nO);
  for (int g = 0; g < cnt; ++g) node(g, rv);
  storePadded<ROWS, kBlock>(pend, res, waveBase, padT);
}

// Lane-permuted copy of a double within a wave through DPP moves (VALU, no
// LDS traffic). CTRL is the DPP control: quad_perm (0x00-0xff), row_shr:n
// (0x110 + n), row_bcast:15 (0x142), row_bcast:31 (0x143). With every row
// enabled, lanes whose source lies outside their row read 0 (the bound
// control bit), so no zeroed register has to be prepared; lanes in rows
// outside ROW_MASK are not written and keep the 0 they start from.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dppD(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  int lo, hi;
  if constexpr (ROW_MASK == 0xf) {
    lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xf, 0xf, true);
    hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, true);
  } else {
    lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROW_MASK, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROW_MASK, 0xf, false);
  }
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}

// Inclusive prefix sum over the 64 lanes of a wave (f64, fixed order):
// within each row of 16 lanes by shifts of 1, 2, 4, 8, then row 0's total
// into row 1 and row 2's into row 3, then lane 31's into rows 2 and 3.
__device__ __forceinline__ double waveInclusiveSum(double x) {
  x += dppD<0x111>(x);
  x += dppD<0x112>(x);
  x += dppD<0x114>(x);
  x += dppD<0x118>(x);
  x += dppD<0x142, 0xa>(x);
  x += dppD<0x143, 0xc>(x);
  return x;
}

// LDS of a fold group: the staged reach row and its zero sentinel slot nO.
__host__ __device__ constexpr size_t foldLdsBytes(int nO) { return (size_t)(((nO + 1) * 4 + 15) & ~15); }

// LDS of a showdown group: reach (+ sentinel) | strength-order prefixes
// (f64, 4 per lane) | per-card segment prefixes (f32, one extra slot per
// card holding its total).
__host__ __device__ constexpr size_t sdRawBytes(int nO) { return (size_t)(((nO + 1) * 4 + 15) & ~15); }
__host__ __device__ constexpr size_t sdLdsBytes(int nO, int lanes) {
  return sdRawBytes(nO) + (size_t)lanes * 4 * sizeof(double) + (size_t)(2 * nO + 52) * 4;
}

// Showdown group. Per node: opponent reach is staged in LDS, then
//  - the class's opponent slots in strength order (4 per lane) are gathered
//    and scanned in f64 into the prefixes Xs (entry-major: position p at
//    Xs[(p % 4) * NT + p / 4], so a wave's stores of one entry are
//    consecutive);
//  - each card's segment (its opponent slots in strength order) is gathered
//    by 4 lanes, and the segment-relative prefixes are stored in f32 (Y,
//    segment c at offset c past its position so each segment has a slot
//    for its total); segments hold at most 48 slots, so their f32 prefixes
//    keep about 2^-24 of a segment total, not of the whole range.
// Per traverser combo the weaker / weaker-or-equal / all masses are
//   win = Xs[le] - Y1[a1] - Y2[a2], leq = Xs[te] - Y1[b1] - Y2[b2] + same,
//   all = total - t1 - t2 + same
// with the compile-time cut points le, te (strength list) and a, b (card
// segments). The next node's reach row is requested as soon as this one is
// staged.
template <int NT, int ROWS>
__device__ void showdownGroup(const Dev& d, const Item& it, char* lds) {
  __shared__ double waveTot[NT / 64];
  __shared__ float segTot[52];
  __shared__ TermNode grp[kMaxGroup];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar tests
  const int nO = it.m, nT = it.n, cnt = it.cnt;
  const int waveBase = wave * 64, padT = lineEnd(nT);
  float* raw = reinterpret_cast<float*>(lds);
  double* Xs = reinterpret_cast<double*>(lds + sdRawBytes(nO));
  float* Y = reinterpret_cast<float*>(lds + sdRawBytes(nO) + (size_t)NT * 4 * sizeof(double));
  const TermNode* nodes = stageNodes(grp, at<TermNode>(d, it.a), cnt);
  const char* base = d.arena + it.b;

  // class data for the whole group: 4 strength-order slots, this lane's part
  // of one card segment, the per-combo bounds. Every load is issued by every
  // lane (indices clamped into the tables) and the lanes that need no value
  // drop it: a load under a per-lane condition is compiled to a branch with
  // its own wait, one round trip after another
  const uint2 gs = at<uint2>(d, it.c + kGatherHeader)[tid];
  const int sg = tid >> 2, part = tid & 3;  // lane groups 52.. have no segment
  // the card of this lane group (the groups take the cards longest segment first)
  const int sc = at<uint8_t>(d, it.c + kCardOrderAt)[min(sg, 51)];
  int segLo, segLen;
  {
    const int32_t* so = at<int32_t>(d, it.c);
    const int s0 = so[sc], sl = so[sc + 1] - s0;
    segLo = sg < 52 ? s0 + (sl * part) / 4 : 0;
    segLen = sg < 52 ? s0 + (sl * (part + 1)) / 4 - segLo : 0;
  }
  uint16_t idx[kFoldPart];
  {
    const uint16_t* seg = at<uint16_t>(d, it.c + kGatherHeader + (int64_t)NT * 4 * 2);
    const int last = max(2 * nO - 1, 0);  // the per-card lists hold 2 nO slots
#pragma unroll
    for (int k = 0; k < kFoldPart; ++k) {
      const uint16_t e = seg[min(segLo + k, last)];
      idx[k] = k < segLen ? e : (uint16_t)nO;
    }
  }
  const int partMax = waveMax(segLen);  // the wave's longest part
  // the wave's strength-order positions (256 w .. 256 w + 255) reach the
  // list's end nO: waves past it have nothing to gather or scan
  const bool scanWave = 4 * waveBase <= nO;
  uint4 meta[ROWS];
  {
    const uint4* md = at<uint4>(d, it.d);
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      const int i = tid + q * NT;
      const uint4 v = md[min(i, max(nT - 1, 0))];
      meta[q] = i < nT ? v : make_uint4(0, 0, 0, 0);
    }
  }

  float res[ROWS] = {};  // lanes past the list keep 0 (stored into the row's padding)
  float* pend = nullptr;  // the previous node's output row
  auto node = [&](int g, float(&rv)[ROWS]) {
    const TermNode nd = nodes[g];
    if (g > 0) __syncthreads();  // the previous node's queries are done with raw / Xs / Y
    // 1. stage reach (raw[nO] = 0 is the padding sentinel)
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      const int j = tid + q * NT;
      if (j < nO) raw[j] = rv[q];
    }
    if (tid == 0) raw[nO] = 0.0f;
    if (g > 0) storePadded<ROWS, NT>(pend, res, waveBase, padT);
    __syncthreads();
    loadReach<ROWS, NT>(rv, base, nodes, g + 1, cnt, nO);  // the next node's reach

    // 2. gathers; strength-order scan across the block
    float gv[4] = {};
    double own = 0.0, inc = 0.0;
    if (scanWave) {
      gv[0] = raw[gs.x & 0xffffu];
      gv[1] = raw[gs.x >> 16];
      gv[2] = raw[gs.y & 0xffffu];
      gv[3] = raw[gs.y >> 16];
      own = (((double)gv[0] + (double)gv[1]) + (double)gv[2]) + (double)gv[3];
      inc = waveInclusiveSum(own);
      if (lane == 63) waveTot[wave] = inc;
    }
    float sv[kFoldPart];
    gatherPart(sv, raw, idx, partMax);
    double segSum = 0.0;
#pragma unroll
    for (int k = 0; k < kFoldPart; ++k) {
      if (k % kPartChunk == 0 && k >= partMax) break;
      segSum += (double)sv[k];
    }
    __syncthreads();
    if (scanWave) {  // the waves below a scanning wave scan too
      double run = inc - own;
#pragma unroll
      for (int w = 0; w < NT / 64 - 1; ++w) {
        const double t = waveTot[w];
        run += w < wave ? t : 0.0;  // +0.0 past this wave: the same sum, no loop
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        Xs[k * NT + tid] = run;
        run += (double)gv[k];
      }
    }
    // 3. segment-relative prefixes: the quad's four parts, in order
    const double p0 = dppD<0x00>(segSum), p1 = dppD<0x55>(segSum);
    const double p2 = dppD<0xaa>(segSum), p3 = dppD<0xff>(segSum);
    if (sg < 52) {
      double y = part == 0 ? 0.0 : part == 1 ? p0 : part == 2 ? p0 + p1 : (p0 + p1) + p2;
      float* yo = Y + segLo + sc;
#pragma unroll
      for (int k = 0; k < kFoldPart; ++k) {
        if (k % kPartChunk == 0 && k >= partMax) break;
        if (k < segLen) yo[k] = (float)y;
        y += (double)sv[k];
      }
      if (part == 3) {
        const float t = (float)(((p0 + p1) + p2) + p3);
        yo[segLen] = t;
        segTot[sc] = t;
      }
    }
    __syncthreads();

    // 4. per traverser combo
    auto xs = [&](int p) -> double { return Xs[(p & 3) * NT + (p >> 2)]; };
    pend = reinterpret_cast<float*>(d.arena + it.b + nd.out);
    // value = vWin win + vTie (leq - win) + vLose (all - leq)
    //       = kWin win + kLeq leq + vLose all
    const double vLose = nd.v2;
    const double kWin = (double)nd.v0 - (double)nd.v1, kLeq = (double)nd.v1 - vLose;
    const double total = xs(nO);
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      if ((int)tid + q * NT >= nT) continue;  // whole waves past the list skip
      const uint4 m = meta[q];
      const int le = m.x & 0xffff, te = m.x >> 16;
      const int c1 = m.w & 0xff, c2 = (m.w >> 8) & 0xff;
      const int a1 = (m.y & 0xffff) + c1, b1 = (m.y >> 16) + c1;
      const int a2 = (m.z & 0xffff) + c2, b2 = (m.z >> 16) + c2;
      const int same = (int)(m.w >> 16) - 1;
      // the nine lookups are issued together (no branch on `same`: the
      // sentinel raw[nO] is 0), then the arithmetic
      const float fs = raw[same >= 0 ? same : nO];
      const double xl = xs(le), xt = xs(te);
      const float y1 = Y[a1], y2 = Y[a2], y3 = Y[b1], y4 = Y[b2];
      const float t1 = segTot[c1], t2 = segTot[c2];
      __builtin_amdgcn_sched_group_barrier(0x0100, 9, 0);  // LDS reads
      __builtin_amdgcn_sched_group_barrier(0x0002, 64, 0);  // then VALU
      // each card pair is summed in f32 first: the sum is commutative, so a
      // combo and its suit-swapped twin get the same bits whichever card
      // comes first; masses below 0 (rounding) count as 0
      const double rs = (double)fs;
      const double win = fmax(xl - (double)(y1 + y2), 0.0);
      const double leq = fmax(xt - (double)(y3 + y4) + rs, 0.0);
      const double all = fmax(total - (double)(t1 + t2) + rs, 0.0);
      res[q] = (float)(kWin * win + (kLeq * leq + vLose * all));
    }
  };
  float rv[ROWS];
  loadReach<ROWS, NT>(rv, base, nodes, 0, cnt, nO);
  for (int g = 0; g < cnt; ++g) node(g, rv);
  storePadded<ROWS, NT>(pend, res, waveBase, padT);
}

// Terminal launches of a half-step (after the forward pass has produced all
// reach rows): one for the fold groups, one for the showdown groups, so each
// kernel keeps its own register budget.
// masked: some listed spot is retired; skip its items.
__global__ __launch_bounds__(kBlock) void foldKernel(Dev d, const Item* __restrict__ items,
                                                     int masked) {
  extern __shared__ __align__(16) char lds[];
  const Item it = items[blockIdx.x];
  if (masked && !d.ctl[it.f].alive) return;
  if (it.na == 3) foldGroup<3>(d, it, lds);
  else foldGroup<kRowRegs>(d, it, lds);
}

// Showdown groups of classes with at most 768 combos per side: 256 lanes,
// 3 combos and 4 strength-order slots each. Held to 96 VGPRs for 5 waves
// per SIMD instead of 4 (two spilled; 3 % faster on the 256-spot batch).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5, 8))) void showdownKernel(Dev d, const Item* __restrict__ items,
                                                         int masked) {
  extern __shared__ __align__(16) char lds[];
  const Item it = items[blockIdx.x];
  if (masked && !d.ctl[it.f].alive) return;
  showdownGroup<kBlock, 3>(d, it, lds);
}

// The larger classes (up to 1326 combos per side): 512 lanes.
__global__ __launch_bounds__(kWideBlock) void showdownWideKernel(Dev d,
                                                                 const Item* __restrict__ items,
                                                                 int masked) {
  extern __shared__ __align__(16) char lds[];
  const Item it = items[blockIdx.x];
  if (masked && !d.ctl[it.f].alive) return;
  showdownGroup<kWideBlock, 3>(d, it, lds);
}

// Forward launch of one depth: opponent reach of the next depth.
// kindMask: profiling attribution only (-1 = every kind in real launches).
template <int RD, int RC>
__global__ __launch_bounds__(kBlock) void fwdKernel(Dev d, const Item* __restrict__ items,
                                                    int count, int fromStrat, int kindMask,
                                                    int masked) {
  const int w = waveItem();
  const Item it = fetchItem(d, items, w, count);
  if (w >= count) return;
  if (!((kindMask >> it.kind) & 1)) return;
  if (masked && !d.ctl[it.f].alive) return;
  fwdItem<RD, RC>(d, it, fromStrat);
}

// ---------------------------------------------------------------------------
// Backward.
// ---------------------------------------------------------------------------

// Decide tile: every lane handles up to R traverser combos kWave apart
// (all row loads issued first). NA > 0: action count fixed at compile time
// (values stay in registers); NA == 0: generic loop.
// dry (CFR mode, timing only): the update is computed and stored, but the
// stored values are the old ones, so the tile moves exactly the bytes of a
// real update and changes no row.
// Whole lines, as in the forward pass (lineEnd): the lanes past the n
// traverser combos in the row's last line work on its padding.
template <int MODE, int NA, int R>
__device__ __forceinline__ void decideTile(const Dev& d, const Item& it, int dry) {
  const int c0 = it.tile * kWave * R + waveLane();
  const int n = lineEnd(it.n);
  if (c0 >= n) return;
  const int na = NA > 0 ? NA : it.na;
  const int64_t stride = it.m, ld = it.ld;
  // loads of every lane, index clamped into the row's last line
  auto cl = [n](int c) { return min(c, n - 1); };
  const float* __restrict__ kid = at<float>(d, it.c);
  float* __restrict__ out = at<float>(d, it.b);
  if (!it.e || MODE == PPSG_MODE_BR) {
    // opponent acts: the child values add up (in action order); best
    // response: their maximum. With the action count fixed at compile time
    // every child load is issued before the first add.
    const bool sum = !it.e;
    float v[R];
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = sum ? 0.0f : -1e30f;
    if (NA > 0) {
      constexpr int A = NA > 0 ? NA : 1;
      float x[R][A];
#pragma unroll
      for (int k = 0; k < R; ++k)
#pragma unroll
        for (int a = 0; a < A; ++a)
          x[k][a] = kid[a * stride + cl(c0 + k * kWave)];
#pragma unroll
      for (int k = 0; k < R; ++k)
#pragma unroll
        for (int a = 0; a < A; ++a) v[k] = sum ? v[k] + x[k][a] : fmaxf(v[k], x[k][a]);
    } else {
      for (int a = 0; a < na; ++a)
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const int c = c0 + k * kWave;
          if (c < n) {
            const float x = kid[a * stride + c];
            v[k] = sum ? v[k] + x : fmaxf(v[k], x);
          }
        }
    }
#pragma unroll
    for (int k = 0; k < R; ++k)
      if (c0 + k * kWave < n) out[c0 + k * kWave] = v[k];
    return;
  }
  constexpr bool plus = MODE == PPSG_MODE_CFR;
  float* __restrict__ src = at<float>(d, plus ? it.a : it.d);
  float* __restrict__ st = at<float>(d, it.d);
  const float uniform = 1.0f / (float)na;
  if (NA > 0) {
    constexpr int A = NA > 0 ? NA : 1;
    float x[R][A], v[R][A], sm[R][A];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int c = c0 + k * kWave;
#pragma unroll
      for (int a = 0; a < A; ++a) {
        x[k][a] = src[a * ld + cl(c)];
        v[k][a] = kid[a * stride + cl(c)];
        sm[k][a] = plus ? st[a * ld + cl(c)] : 0.0f;
      }
    }
    float pos = 0.0f, neg = 0.0f, avg = 0.0f;
    if (MODE == PPSG_MODE_CFR) {
      const Ctl& ctl = d.ctl[it.f];
      pos = ctl.cur[0];
      neg = ctl.cur[1];
      avg = ctl.cur[2];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int c = c0 + k * kWave;
      if (c >= n) continue;
      float acc = 0.0f;
#pragma unroll
      for (int a = 0; a < A; ++a) acc += plus ? (x[k][a] > 0.0f ? x[k][a] : 0.0f) : x[k][a];
      float s[A], val = 0.0f;
#pragma unroll
      for (int a = 0; a < A; ++a) {
        s[a] = sigmaOf(x[k][a], acc, uniform, plus);
        val += s[a] * v[k][a];
      }
      out[c] = val;
      if (MODE != PPSG_MODE_CFR) continue;
#pragma unroll
      for (int a = 0; a < A; ++a) {
        const float ns = sm[k][a] * avg + s[a];
        const float nr = x[k][a] * (x[k][a] >= 0.0f ? pos : neg) + (v[k][a] - val);
        st[a * ld + c] = dry ? sm[k][a] : ns;
        src[a * ld + c] = dry ? x[k][a] : nr;
      }
    }
    return;
  }
  for (int k = 0; k < R; ++k) {
    const int c = c0 + k * kWave;
    if (c >= n) return;
    float acc = 0.0f;
    for (int a = 0; a < na; ++a) {
      const float x = src[a * ld + c];
      acc += plus ? (x > 0.0f ? x : 0.0f) : x;
    }
    float val = 0.0f;
    for (int a = 0; a < na; ++a)
      val += sigmaOf(src[a * ld + c], acc, uniform, plus) * kid[a * stride + c];
    out[c] = val;
    if (MODE != PPSG_MODE_CFR) continue;
    const Ctl& ctl = d.ctl[it.f];
    const float pos = ctl.cur[0], neg = ctl.cur[1], avg = ctl.cur[2];
    for (int a = 0; a < na; ++a) {
      const float x = src[a * ld + c];
      const float sg = sigmaOf(x, acc, uniform, plus);
      const float so = st[a * ld + c];
      st[a * ld + c] = dry ? so : so * avg + sg;
      src[a * ld + c] = dry ? x : x * (x >= 0.0f ? pos : neg) + (kid[a * stride + c] - val);
    }
  }
}

// Chance tile, one lane per traverser (parent) combo: sum of the child
// values over the branches the combo survives, in branch order, accumulated
// in f64 and rounded to f32 once (postflop-solver's chance sum,
// solver.rs:191-199). Combos that a suit swap maps onto each other get the
// same terms in different orders; the f64 sum makes their values equal
// bit for bit, which a suit-isomorphic tree relies on (f32 rounding
// differences there grow under regret matching and break the symmetry the
// deal tables assume). Branches go in chunks: all map loads of a chunk,
// then all value loads, then the adds. Every load is issued whatever its
// lane needs (indices clamped into range) and the conditions are applied
// to the loaded values: a load under a per-lane condition is compiled to
// its own branch with a wait after it, which made each lane's 48 value
// loads one dependent round trip after another (29 us for a launch with a
// chance node of a spot alone, 9 us now). A dropped term adds +0.0, which
// leaves the sum's bits as they are (a sum that starts at +0.0 never
// becomes -0.0). it.e: the deal table serves every card by its own child
// (no suit isomorphism), so no child index is loaded.
__device__ __forceinline__ void chanceTile(const Dev& d, const Item& it) {
  const int i = it.tile * kWave + waveLane();
  const int n = it.n;
  if (i >= lineEnd(n)) return;  // whole lines: lanes past n store 0
  const bool in = i < n;
  const uint32_t ic = (uint32_t)min(i, n - 1);  // lanes past n read slot n-1 (then drop it)
  // 32-bit unsigned byte offsets from uniform bases (a branch block is at
  // most 255 x 1326 floats): one VGPR per address instead of two
  const char* __restrict__ map = at<char>(d, it.a);
  const char* __restrict__ kid = at<char>(d, it.c);
  const uint8_t* __restrict__ sib = at<uint8_t>(d, it.d);  // child serving each dealt card
  const uint32_t stride = (uint32_t)it.m;
  const int nb = it.na;
  const bool ident = it.e != 0;
  constexpr int kChunk = 16;
  double v = 0.0;
  for (int b0 = 0; b0 < nb; b0 += kChunk) {
    int slot[kChunk];
    uint32_t child[kChunk];
    float x[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      const uint32_t b = (uint32_t)min(b0 + k, nb - 1);
      child[k] = ident ? b : (uint32_t)sib[b];
      slot[k] = *reinterpret_cast<const int16_t*>(map + 2u * (b * (uint32_t)n + ic));
    }
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      x[k] = *reinterpret_cast<const float*>(kid + 4u * (child[k] * stride + (uint32_t)max(slot[k], 0)));
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      v += (double)(in && b0 + k < nb && slot[k] >= 0 ? x[k] : 0.0f);
  }
  at<float>(d, it.b)[i] = (float)v;
}

template <int MODE, int RB>
__global__ __launch_bounds__(kBlock) void bwdKernel(Dev d, const Item* __restrict__ items,
                                                    int count, int kindMask, int masked, int dry) {
  const int w = waveItem();
  const Item it = fetchItem(d, items, w, count);
  if (w >= count) return;
  if (!((kindMask >> it.kind) & 1)) return;
  if (masked && !d.ctl[it.f].alive) return;
  if (it.kind == kBwdChance) {
    chanceTile(d, it);
    return;
  }
  switch (it.na) {
    case 1: decideTile<MODE, 1, RB>(d, it, dry); break;
    case 2: decideTile<MODE, 2, RB>(d, it, dry); break;
    case 3: decideTile<MODE, 3, RB>(d, it, dry); break;
    case 4: decideTile<MODE, 4, RB>(d, it, dry); break;
    case 5: decideTile<MODE, 5, RB>(d, it, dry); break;
    case 6: decideTile<MODE, 6, RB>(d, it, dry); break;
    default: decideTile<MODE, 0, RB>(d, it, dry); break;
  }
}

// Bumps the schedule index of every active spot (the last node of an
// iteration's graph) and fetches the next iteration's coefficients (runs
// never pass the table's end, and the entry there is zero).
__global__ void advanceCounters(Ctl* ctl, const int* __restrict__ slots, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    Ctl& c = ctl[slots[i]];
    const long long t = c.counter + 1;
    c.counter = t;
    if (c.coef)
      for (int k = 0; k < 3; ++k) c.cur[k] = c.coef[3 * t + k];
  }
}

// Coefficients of a spot's current schedule index (after the host sets the
// index or the table); zeros outside the table's len entries.
__global__ void refreshCoef(Ctl* ctl, int slot, long long len) {
  if (threadIdx.x != 0) return;
  Ctl& c = ctl[slot];
  const bool in = c.coef && c.counter >= 0 && c.counter < len;
  for (int k = 0; k < 3; ++k) c.cur[k] = in ? c.coef[3 * c.counter + k] : 0.0f;
}

// Sets every control block's alive flag (walks of a subset of spots mask all
// others, then restore the flags of the live set).
__global__ void setAliveFlags(Ctl* ctl, const int* __restrict__ flags, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    ctl[i].alive = flags[i];
}

// ---------------------------------------------------------------------------
// Work-list compaction on the device: a stable filter keeping the items of
// live spots (control block alive flag), in three steps: per-block counts,
// one exclusive scan of the counts, scatter. Each block takes kCompactItems
// consecutive items, every lane four of them in order.
// ---------------------------------------------------------------------------

constexpr int kCompactItems = 4 * kBlock;

__device__ __forceinline__ int blockExclusiveScan(int v, int* total) {
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  int base = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wave) base += wsum[w];
    all += wsum[w];
  }
  __syncthreads();
  if (total) *total = all;
  return base + inc - v;
}

// One list of a compaction pass: its items go from `in` to `out` (the
// list's spare buffer); its blocks are [block0, block0 + nb) of the grid.
struct CompactJob {
  const Item* in;
  Item* out;
  int64_t n;
  int64_t block0;
  int64_t nb;
};

// The job owning grid block b (jobs sorted by block0).
__device__ __forceinline__ int jobOf(const CompactJob* __restrict__ jobs, int nJobs, int64_t b) {
  int lo = 0, hi = nJobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void compactCount(const CompactJob* __restrict__ jobs,
                                                       int nJobs, const Ctl* __restrict__ ctl,
                                                       int* __restrict__ counts) {
  const CompactJob J = jobs[jobOf(jobs, nJobs, blockIdx.x)];
  const int64_t i0 = ((int64_t)blockIdx.x - J.block0) * kCompactItems + 4 * (int64_t)threadIdx.x;
  int c = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (i0 + k < J.n && ctl[J.in[i0 + k].f].alive) ++c;
  int total = 0;
  (void)blockExclusiveScan(c, &total);
  if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// One block per job: its block counts -> exclusive offsets (in place);
// totals[job] = its kept items.
__global__ __launch_bounds__(kBlock) void compactScan(const CompactJob* __restrict__ jobs,
                                                      int* __restrict__ counts,
                                                      int* __restrict__ totals) {
  const CompactJob J = jobs[blockIdx.x];
  int* c = counts + J.block0;
  const int nb = (int)J.nb;
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kBlock) {
    const int i = b0 + (int)threadIdx.x;
    const int v = i < nb ? c[i] : 0;
    int sum = 0;
    const int ex = blockExclusiveScan(v, &sum);
    if (i < nb) c[i] = carry + ex;
    carry += sum;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ __launch_bounds__(kBlock) void compactScatter(const CompactJob* __restrict__ jobs,
                                                         int nJobs, const Ctl* __restrict__ ctl,
                                                         const int* __restrict__ offs) {
  const CompactJob J = jobs[jobOf(jobs, nJobs, blockIdx.x)];
  const int64_t i0 = ((int64_t)blockIdx.x - J.block0) * kCompactItems + 4 * (int64_t)threadIdx.x;
  bool keep[4];
  int c = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    keep[k] = i0 + k < J.n && ctl[J.in[i0 + k].f].alive;
    c += keep[k];
  }
  int pos = offs[blockIdx.x] + blockExclusiveScan(c, nullptr);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (keep[k]) J.out[pos++] = J.in[i0 + k];
}

// Plain streaming copy (16 B per lane per step): the calibration kernel for
// hardware byte counters.
// One float4 per lane, one 4 KiB tile per block (the fastest plain copy
// shape measured here; a grid-stride loop over 4096 blocks is ~25 % slower).
__global__ __launch_bounds__(kBlock) void copyProbe(const float4* __restrict__ a,
                                                    float4* __restrict__ b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// ---------------------------------------------------------------------------
// Host: arena allocator (first fit over an offset-ordered free list; freed
// blocks merge with free neighbours).
// ---------------------------------------------------------------------------

int64_t roundUp(int64_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

struct Arena {
  char* base = nullptr;
  int64_t cap = 0;
  std::map<int64_t, int64_t> holes;  // offset -> bytes

  int64_t alloc(int64_t bytes) {
    bytes = roundUp(bytes);
    for (auto it = holes.begin(); it != holes.end(); ++it) {
      if (it->second < bytes) continue;
      const int64_t off = it->first, rest = it->second - bytes;
      holes.erase(it);
      if (rest > 0) holes[off + bytes] = rest;
      return off;
    }
    return -1;
  }
  void release(int64_t off, int64_t bytes) {
    bytes = roundUp(bytes);
    auto it = holes.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != holes.end() && it->first + it->second == next->first) {
      it->second += next->second;
      holes.erase(next);
    }
    if (it != holes.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        holes.erase(it);
      }
    }
  }
  int64_t largestHole() const {
    int64_t m = 0;
    for (const auto& h : holes) m = std::max(m, h.second);
    return m;
  }
  int64_t freeBytes() const {
    int64_t m = 0;
    for (const auto& h : holes) m += h.second;
    return m;
  }
};

// Byte layout of one spot's chunk.
struct Layout {
  int64_t bytes = 0;
  int64_t toParent[2] = {0, 0}, bmap[2] = {0, 0}, gather[2] = {0, 0}, dealSib = 0;
  int64_t cardSame[2] = {0, 0}, sdMeta[2] = {0, 0};
  int64_t regret = 0, strat = 0, cfv = 0, reach[2] = {0, 0};
  int64_t weight[2] = {0, 0};  // f64 base weights (root value sums)
  int64_t termNodes = 0;  // TermNode entries of every terminal group
  int64_t add(int64_t b) {
    const int64_t off = bytes;
    bytes += roundUp(std::max<int64_t>(b, 1));
    return off;
  }
};

// Showdown group shape: the 256-lane kernel takes classes with at most
// 3 * kBlock combos per side; the rest go to the 512-lane kernel.
bool wideShowdown(int nT, int nO) { return std::max(nT, nO) > 3 * kBlock; }

// Dynamic LDS of a terminal item: reach (+ sentinel), and for showdown the
// prefixes.
size_t terminalLdsOf(const Item& it, int lanes) {
  return it.kind == kBwdShowdown ? sdLdsBytes(it.m, lanes) : foldLdsBytes(it.m);
}

// Host-side tables of one spot, built from its description.
struct SpotTables {
  std::vector<char> gather[2];  // per class and terminal kind, see gatherBlock
  std::vector<int64_t> sdOff[2], foldOff[2];  // class -> byte offset, -1: none
  std::vector<uint32_t> cardSame[2];  // per list entry: c1 | c2 << 8 | (same+1) << 16
  std::vector<TermNode> termNodes;    // grouped: each group's nodes are contiguous

  // showdown: [53 i32 segment bounds | pad to 256 B | u16 slots in strength
  // order, padded with the sentinel slot n to 4 per lane of the class's
  // kernel | u16 per-card segments]; fold: [53 i32 segment bounds | pad |
  // u16 per-card segments]
  void gatherBlock(const ppsg_tree_desc* t, int p, int cl, bool sd) {
    auto& offs = sd ? sdOff[p] : foldOff[p];
    if (offs[cl] >= 0) return;
    const int n = t->cls_n[p][cl];
    const int64_t lo = t->cls_list_off[p][cl];
    std::vector<uint16_t> list;
    if (sd) {
      const bool wide = wideShowdown(t->cls_n[0][cl], t->cls_n[1][cl]);
      list.assign(t->sorted[p] + lo, t->sorted[p] + lo + n);
      list.resize((size_t)4 * (wide ? kWideBlock : kBlock), (uint16_t)n);
    }
    list.insert(list.end(), t->seg_slot[p] + 2 * lo, t->seg_slot[p] + 2 * lo + 2 * n);
    std::vector<char>& g = gather[p];
    g.resize((g.size() + 15) / 16 * 16, 0);
    offs[cl] = (int64_t)g.size();
    const size_t at0 = g.size();
    g.resize(at0 + kGatherHeader + list.size() * 2, 0);
    const int32_t* so = t->seg_off[p] + (int64_t)cl * 53;
    std::memcpy(g.data() + at0, so, 53 * 4);
    // lane group k (lanes 4k..4k+3) takes card order[k]: longest segments
    // first, so the 16 cards of a wave have similar segment lengths and a
    // wave's longest part (which sets its gather chunks) is close to the rest
    uint8_t order[52];
    for (int c = 0; c < 52; ++c) order[c] = (uint8_t)c;
    std::stable_sort(order, order + 52, [&](uint8_t a, uint8_t b) {
      return so[a + 1] - so[a] > so[b + 1] - so[b];
    });
    std::memcpy(g.data() + at0 + kCardOrderAt, order, 52);
    std::memcpy(g.data() + at0 + kGatherHeader, list.data(), list.size() * 2);
  }
};

enum SpotState { kFree = 0, kLive = 1, kParked = 2, kDoomed = 3 };

// The fold (or showdown) nodes of one board class for one traverser; the
// engine cuts it into work items of up to its group size.
struct TermGroup {
  uint8_t kind = 0;
  uint8_t rows = 0;
  bool wide = false;  // showdown: the 512-lane kernel
  int nT = 0, nO = 0, count = 0;
  int64_t first = 0;          // index of its first TermNode
  int64_t gather = 0, meta = 0;  // chunk-relative
};

struct Spot {
  int state = kFree;
  bool inLists = false;         // its items are in the merged work lists
  int64_t off = -1, bytes = 0;  // arena chunk
  float* coefDev = nullptr;     // schedule table (own allocation)
  int64_t coefCap = 0, coefLen = 0;
  int64_t counter = 0;          // host mirror of the device schedule index
  int n[2] = {0, 0};
  int maxDepth = 0;
  Layout L;
  int64_t regTotal = 0, cfvTotal = 0, reachTotal[2] = {0, 0};
  int64_t rootCfv = 0;  // chunk-relative byte offset of the root cfv row
  // work items per [traverser][0 forward, 1 backward][depth], chunk-relative,
  // and the node each item serves (subtree walks select by node)
  std::vector<std::vector<Item>> items[2][2];
  std::vector<std::vector<int32_t>> itemNode[2][2];
  std::vector<int32_t> termNode;  // node of each terminal entry (group order)
  // depth of each terminal entry; a group's entries run by depth, then node
  std::vector<int16_t> termDepth;
  int64_t nodes = 0;
  std::vector<TermGroup> groups[2];  // per traverser
  int group = 4;  // terminal nodes per work item, set when the spot is added
  int64_t itemCount = 0;  // listed items (terminal items at the group size in use)
  double fwdBytes = 0.0, bwdBytes = 0.0;
  double kindBytes[6] = {0, 0, 0, 0, 0, 0};  // per ItemKind, one iteration
};

// Plans a spot: chunk layout, host tables and chunk-relative work items.
int planSpot(const ppsg_tree_desc* t, Spot& S, SpotTables& T) {
  if (!t || t->n_nodes <= 0 || t->n_base[0] <= 0 || t->n_base[1] <= 0) return fail("empty tree");
  if (!t->weight[0] || !t->weight[1]) return fail("tree description lacks the base weights");
  const int nc = t->n_class;
  const int64_t N = t->n_nodes;
  for (int p = 0; p < 2; ++p)
    for (int c = 0; c < nc; ++c)
      if (t->cls_n[p][c] > kMaxList) return fail("combo list longer than 1326");
  auto kidsOf = [&](int64_t u, int& b, int& e) {
    b = t->child_base[u];
    e = u + 1 < N ? t->child_base[u + 1] : (int)t->n_child;
  };
  if (t->max_depth < 0 || t->max_depth > 32767) return fail("tree description: max_depth out of range");
  // rows on 64-float lines; a decide node's action rows PPSG_ROW_STRIDE apart
  for (int64_t u = 0; u < N; ++u) {
    if (t->depth[u] < 0 || t->depth[u] > t->max_depth)
      return fail("tree description: node depth out of range");
    if (t->cfv_off[u] % 64 || t->reach_off[0][u] % 64 || t->reach_off[1][u] % 64)
      return fail("tree description: a row does not start on a 64-float line");
    if (t->kind[u] == PPSG_DECIDE &&
        (t->reg_off[u] % 64 ||
         t->reg_off[u] + t->na[u] * PPSG_ROW_STRIDE(t->cls_n[t->owner[u]][t->cls[u]]) >
             t->reg_total))
      return fail("tree description: regret rows not PPSG_ROW_STRIDE apart within reg_total");
  }
  for (int p = 0; p < 2; ++p) {
    T.sdOff[p].assign(nc, -1);
    T.foldOff[p].assign(nc, -1);
  }
  // every card segment of a class with terminals fits 4 lanes x kFoldPart
  std::vector<uint8_t> termCls(nc, 0);
  for (int64_t u = 0; u < N; ++u)
    if (t->kind[u] == PPSG_SHOWDOWN || t->kind[u] == PPSG_FOLD) termCls[t->cls[u]] = 1;
  for (int c = 0; c < nc; ++c) {
    if (!termCls[c]) continue;
    for (int p = 0; p < 2; ++p) {
      const int32_t* so = t->seg_off[p] + (int64_t)c * 53;
      for (int k = 0; k < 52; ++k)
        if (so[k + 1] - so[k] > 4 * kFoldPart) return fail("terminal: card segment too long");
    }
  }
  for (int64_t u = 0; u < N; ++u) {
    const int cl = t->cls[u];
    if (t->kind[u] == PPSG_CHANCE) {
      int b, e;
      kidsOf(u, b, e);
      if (t->deal_n[cl] > 255 || t->deal_off[cl] < 0 || e <= b)
        return fail("chance node without a valid deal table");
      for (int32_t k = 0; k < t->deal_n[cl]; ++k)
        if (t->deal_sib[t->deal_off[cl] + k] >= e - b) return fail("deal table names a missing child");
    } else if (t->kind[u] == PPSG_SHOWDOWN || t->kind[u] == PPSG_FOLD) {
      const bool sd = t->kind[u] == PPSG_SHOWDOWN;
      for (int p = 0; p < 2; ++p) T.gatherBlock(t, p, cl, sd);
    }
  }
  for (int p = 0; p < 2; ++p) {
    T.cardSame[p].resize(t->list_total[p]);
    for (int64_t e = 0; e < t->list_total[p]; ++e)
      T.cardSame[p][e] =
          (uint32_t)t->list_cards[p][e] | ((uint32_t)(t->list_same[p][e] + 1) << 16);
  }
  Layout& L = S.L;
  for (int p = 0; p < 2; ++p) {
    L.toParent[p] = L.add(2 * t->list_total[p]);
    L.bmap[p] = L.add(2 * t->deal_map_total[p]);
    L.gather[p] = L.add((int64_t)T.gather[p].size());
    L.cardSame[p] = L.add(4 * t->list_total[p]);
    L.sdMeta[p] = L.add(16 * t->list_total[p]);
  }
  L.dealSib = L.add(t->deal_total);
  for (int p = 0; p < 2; ++p) L.weight[p] = L.add(8 * (int64_t)t->n_base[p]);
  L.regret = L.add(4 * t->reg_total);
  L.strat = L.add(4 * t->reg_total);
  L.cfv = L.add(4 * t->cfv_total);
  L.reach[0] = L.add(4 * t->reach_total[0]);
  L.reach[1] = L.add(4 * t->reach_total[1]);
  S.n[0] = t->n_base[0];
  S.n[1] = t->n_base[1];
  S.maxDepth = t->max_depth;
  S.regTotal = t->reg_total;
  S.cfvTotal = t->cfv_total;
  S.reachTotal[0] = t->reach_total[0];
  S.reachTotal[1] = t->reach_total[1];
  S.rootCfv = L.cfv + 4 * t->cfv_off[0];

  const int D = t->max_depth + 1;
  S.nodes = N;
  for (int tr = 0; tr < 2; ++tr)
    for (int ps = 0; ps < 2; ++ps) {
      S.items[tr][ps].assign(D, {});
      S.itemNode[tr][ps].assign(D, {});
    }
  auto tiles = [](int n, int r) { return (n + kWave * r - 1) / (kWave * r); };
  std::vector<int> groupOf[2] = {std::vector<int>((size_t)2 * nc, 0), std::vector<int>((size_t)2 * nc, 0)};  // (class, kind) -> group index + 1
  std::vector<std::vector<TermNode>> members[2];
  std::vector<std::vector<int32_t>> memberNode[2];
  for (int64_t u = 0; u < N; ++u) {
    const int dep = t->depth[u], cl = t->cls[u], k = t->kind[u];
    int kb, ke;
    kidsOf(u, kb, ke);
    for (int tr = 0; tr < 2; ++tr) {
      const int op = 1 - tr;
      const int nT = t->cls_n[tr][cl], nO = t->cls_n[op][cl];
      std::vector<Item>& fwd = S.items[tr][0][dep];
      std::vector<Item>& bwd = S.items[tr][1][dep];
      if (k == PPSG_DECIDE) {
        const int na = t->na[u];
        const int first = t->child_flat[kb];
        if (t->owner[u] == op) {
          const int64_t rs =
              na > 1 ? t->reach_off[tr][t->child_flat[kb + 1]] - t->reach_off[tr][first] : 0;
          for (int q = 0; q < tiles(nO, 1); ++q) {
            Item it{};
            it.kind = kFwdDecide;
            it.r = 1;
            it.a = L.regret + 4 * t->reg_off[u];
            it.d = L.strat + 4 * t->reg_off[u];
            it.b = L.reach[tr] + 4 * t->reach_off[tr][u];
            it.c = L.reach[tr] + 4 * t->reach_off[tr][first];
            it.m = (int32_t)rs;
            it.ld = (int32_t)PPSG_ROW_STRIDE(nO);
            it.n = nO;
            it.na = (uint8_t)na;
            it.tile = (uint8_t)q;
            fwd.push_back(it);
          }
        }
        const int64_t stride = na > 1 ? t->cfv_off[t->child_flat[kb + 1]] - t->cfv_off[first] : 0;
        for (int q = 0; q < tiles(nT, 1); ++q) {
          Item it{};
          it.kind = kBwdDecide;
          it.r = 1;
          it.a = L.regret + 4 * t->reg_off[u];
          it.d = L.strat + 4 * t->reg_off[u];
          it.b = L.cfv + 4 * t->cfv_off[u];
          it.c = L.cfv + 4 * t->cfv_off[first];
          it.m = (int32_t)stride;
          it.ld = (int32_t)PPSG_ROW_STRIDE(t->cls_n[t->owner[u]][cl]);
          it.n = nT;
          it.na = (uint8_t)na;
          it.e = t->owner[u] == tr;
          it.tile = (uint8_t)q;
          bwd.push_back(it);
        }
      } else if (k == PPSG_CHANCE) {
        const int nb = ke - kb;
        const float inv = 1.0f / (float)(52 - t->cls_board_n[cl] - 4);
        for (int b = kb; b < ke; ++b) {
          const int ch = t->child_flat[b];
          const int ccl = t->cls[ch];
          for (int q = 0; q < tiles(t->cls_n[op][ccl], 1); ++q) {
            Item it{};
            it.kind = kFwdChance;
            it.r = 1;
            it.a = L.toParent[op] + 2 * t->cls_list_off[op][ccl];
            it.b = L.reach[tr] + 4 * t->reach_off[tr][u];
            it.c = L.reach[tr] + 4 * t->reach_off[tr][ch];
            it.n = t->cls_n[op][ccl];
            it.v0 = inv;
            it.tile = (uint8_t)q;
            fwd.push_back(it);
          }
        }
        const int first = t->child_flat[kb];
        const int64_t stride = nb > 1 ? t->cfv_off[t->child_flat[kb + 1]] - t->cfv_off[first] : 0;
        int identDeal = 1;  // every dealt card served by its own child
        for (int32_t k = 0; k < t->deal_n[cl]; ++k)
          identDeal &= t->deal_sib[t->deal_off[cl] + k] == k;
        for (int q = 0; q < tiles(nT, 1); ++q) {
          Item it{};
          it.kind = kBwdChance;
          it.r = 1;
          it.a = L.bmap[tr] + 2 * t->deal_map_off[tr][cl];
          it.b = L.cfv + 4 * t->cfv_off[u];
          it.c = L.cfv + 4 * t->cfv_off[first];
          it.d = L.dealSib + t->deal_off[cl];
          it.m = (int32_t)stride;
          it.n = nT;
          it.na = (uint8_t)t->deal_n[cl];
          it.e = identDeal;
          it.tile = (uint8_t)q;
          bwd.push_back(it);
        }
      } else {
        const bool sd = k == PPSG_SHOWDOWN;
        const int key = cl * 2 + (sd ? 1 : 0);
        auto& slot = groupOf[tr][key];
        if (slot == 0) {
          TermGroup g;
          g.kind = sd ? kBwdShowdown : kBwdFold;
          // the opponent's gather block of the class; the traverser's
          // per-combo bounds (showdown) or cards (fold)
          g.gather = L.gather[op] + (sd ? T.sdOff[op][cl] : T.foldOff[op][cl]);
          g.meta = sd ? L.sdMeta[tr] + 16 * t->cls_list_off[tr][cl]
                      : L.cardSame[tr] + 4 * t->cls_list_off[tr][cl];
          g.wide = sd && wideShowdown(nT, nO);
          g.rows = (uint8_t)(sd || std::max(nT, nO) <= 3 * kBlock ? 3 : kRowRegs);
          g.nT = nT;
          g.nO = nO;
          S.groups[tr].push_back(g);
          members[tr].emplace_back();
          memberNode[tr].emplace_back();
          slot = (int)S.groups[tr].size();
        }
        TermNode e{};
        e.reach = L.reach[tr] + 4 * t->reach_off[tr][u];
        e.out = L.cfv + 4 * t->cfv_off[u];
        const float* tv = t->term_val + u * 8 + tr * 4;
        e.v0 = tv[0];
        e.v1 = tv[1];
        e.v2 = tv[2];
        members[tr][slot - 1].push_back(e);
        memberNode[tr][slot - 1].push_back((int32_t)u);
      }
      // the items just added serve node u
      S.itemNode[tr][0][dep].resize(fwd.size(), (int32_t)u);
      S.itemNode[tr][1][dep].resize(bwd.size(), (int32_t)u);
    }
  }
  // terminal node entries, contiguous per group, after the rows; within a
  // group by depth, then node, so the entries at depth <= k are a prefix of
  // it for every k (the split terminal lists)
  for (int tr = 0; tr < 2; ++tr)
    for (size_t g = 0; g < S.groups[tr].size(); ++g) {
      const std::vector<int32_t>& ids = memberNode[tr][g];
      std::vector<int32_t> ord(ids.size());
      for (size_t k = 0; k < ord.size(); ++k) ord[k] = (int32_t)k;
      std::stable_sort(ord.begin(), ord.end(),
                       [&](int32_t x, int32_t y) { return t->depth[ids[x]] < t->depth[ids[y]]; });
      S.groups[tr][g].first = (int64_t)T.termNodes.size();
      S.groups[tr][g].count = (int)ids.size();
      for (const int32_t k : ord) {
        T.termNodes.push_back(members[tr][g][k]);
        S.termNode.push_back(ids[k]);
        S.termDepth.push_back((int16_t)t->depth[ids[k]]);
      }
    }
  L.termNodes = L.add((int64_t)(T.termNodes.size() * sizeof(TermNode)));
  S.bytes = L.bytes;
  for (int tr = 0; tr < 2; ++tr)
    for (int dep = 0; dep < D; ++dep)
      S.itemCount += (int64_t)(S.items[tr][0][dep].size() + S.items[tr][1][dep].size());
  // Algorithmic bytes of one iteration per node kind: every f32 row element
  // read or written once per use (index tables excluded). Fold/showdown
  // count with the forward pass (their launches precede the backward one).
  for (int64_t u = 0; u < N; ++u) {
    const int cl = t->cls[u];
    const double n0 = t->cls_n[0][cl], n1 = t->cls_n[1][cl];
    if (t->kind[u] == PPSG_DECIDE) {
      const double a = t->na[u];
      const double np = t->owner[u] == 0 ? n0 : n1, nq = t->owner[u] == 0 ? n1 : n0;
      S.kindBytes[kFwdDecide] += 4 * np * (2 * a + 1);
      S.kindBytes[kBwdDecide] += 4 * np * (5 * a + 1) + 4 * nq * (a + 1);
    } else if (t->kind[u] == PPSG_CHANCE) {
      int b, e;
      kidsOf(u, b, e);
      for (int k = b; k < e; ++k) {
        const int ccl = t->cls[t->child_flat[k]];
        S.kindBytes[kFwdChance] += 8 * ((double)t->cls_n[0][ccl] + t->cls_n[1][ccl]);
      }
      for (int32_t k = 0; k < t->deal_n[cl]; ++k) {
        const int ccl = t->cls[t->child_flat[b + t->deal_sib[t->deal_off[cl] + k]]];
        S.kindBytes[kBwdChance] += 4 * ((double)t->cls_n[0][ccl] + t->cls_n[1][ccl]);
      }
      S.kindBytes[kBwdChance] += 4 * (n0 + n1);
    } else {
      S.kindBytes[t->kind[u] == PPSG_FOLD ? kBwdFold : kBwdShowdown] += 8 * (n0 + n1);
    }
  }
  S.fwdBytes = S.kindBytes[kFwdDecide] + S.kindBytes[kFwdChance] + S.kindBytes[kBwdFold] +
               S.kindBytes[kBwdShowdown];
  S.bwdBytes = S.kindBytes[kBwdDecide] + S.kindBytes[kBwdChance];
  return 0;
}

// One merged work list of a (traverser, pass, depth): the items of every
// listed spot, appended as spots arrive.
struct DevList {
  Item* buf = nullptr;
  int64_t count = 0, cap = 0;
  Item* spare = nullptr;  // compaction target, swapped with buf afterwards
  int64_t spareCap = 0;
  size_t lds = 0;  // forward lists: dynamic LDS of the largest terminal item
};

}  // namespace

// ---------------------------------------------------------------------------
// Launches at a given shape (combos per lane): forward decide RD in {1, 2,
// 4}, forward chance RC in {1, 2, 4, 8}, backward decide RB in {1, 2}.
// ---------------------------------------------------------------------------

namespace {

template <int RD>
void launchFwdRC(int rc, unsigned g, hipStream_t s, Dev d, const Item* buf, int count,
                 int fromStrat, int kindMask, int masked) {
  switch (rc) {
    case 1: fwdKernel<RD, 1><<<g, kBlock, 0, s>>>(d, buf, count, fromStrat, kindMask, masked); break;
    case 2: fwdKernel<RD, 2><<<g, kBlock, 0, s>>>(d, buf, count, fromStrat, kindMask, masked); break;
    case 8: fwdKernel<RD, 8><<<g, kBlock, 0, s>>>(d, buf, count, fromStrat, kindMask, masked); break;
    default: fwdKernel<RD, 4><<<g, kBlock, 0, s>>>(d, buf, count, fromStrat, kindMask, masked); break;
  }
}

void launchFwdShape(int rd, int rc, hipStream_t s, Dev d, const Item* buf, int64_t count,
                    int fromStrat, int kindMask, int masked) {
  const unsigned g = (unsigned)((count + kWavesPerBlock - 1) / kWavesPerBlock);
  const int c = (int)count;
  if (rd == 1) launchFwdRC<1>(rc, g, s, d, buf, c, fromStrat, kindMask, masked);
  else if (rd == 4) launchFwdRC<4>(rc, g, s, d, buf, c, fromStrat, kindMask, masked);
  else launchFwdRC<2>(rc, g, s, d, buf, c, fromStrat, kindMask, masked);
}

template <int MODE>
void launchBwdRB(int rb, unsigned g, hipStream_t s, Dev d, const Item* buf, int count,
                 int kindMask, int masked, int dry) {
  if (rb == 2) bwdKernel<MODE, 2><<<g, kBlock, 0, s>>>(d, buf, count, kindMask, masked, dry);
  else bwdKernel<MODE, 1><<<g, kBlock, 0, s>>>(d, buf, count, kindMask, masked, dry);
}

void launchBwdShape(int mode, int rb, hipStream_t s, Dev d, const Item* buf, int64_t count,
                    int kindMask, int masked, int dry) {
  const unsigned g = (unsigned)((count + kWavesPerBlock - 1) / kWavesPerBlock);
  const int c = (int)count;
  if (mode == PPSG_MODE_CFR) launchBwdRB<PPSG_MODE_CFR>(rb, g, s, d, buf, c, kindMask, masked, dry);
  else if (mode == PPSG_MODE_EV) launchBwdRB<PPSG_MODE_EV>(rb, g, s, d, buf, c, kindMask, masked, 0);
  else launchBwdRB<PPSG_MODE_BR>(rb, g, s, d, buf, c, kindMask, masked, 0);
}

}  // namespace

// ---------------------------------------------------------------------------
// Root value sums: per listed spot, sum_i w[i] * cfv_root[i] over its base
// list for one traverser, in index order, in f64 with separately rounded
// multiplies and adds (the host loop's arithmetic, bit for bit). One 64-lane
// block per spot: the lanes stage 64 entries at a time in LDS, lane 0 adds
// them in order.
// ---------------------------------------------------------------------------

namespace {

struct RootSumJob {
  int64_t cfv, w;  // byte offsets in the arena: root cfv row (f32), weights (f64)
  int32_t n, pad;
};

__global__ __launch_bounds__(kWave) void rootSumKernel(const char* __restrict__ arena,
                                                       const RootSumJob* __restrict__ jobs,
                                                       double* __restrict__ out) {
  __shared__ float c[kWave];
  __shared__ double w[kWave];
  const RootSumJob j = jobs[blockIdx.x];
  const float* __restrict__ cf = reinterpret_cast<const float*>(arena + j.cfv);
  const double* __restrict__ wt = reinterpret_cast<const double*>(arena + j.w);
  const int t = (int)threadIdx.x;
  double s = 0.0;
  for (int b = 0; b < j.n; b += kWave) {
    c[t] = b + t < j.n ? cf[b + t] : 0.0f;
    w[t] = b + t < j.n ? wt[b + t] : 0.0;
    __syncthreads();
    if (t == 0) {
#pragma clang fp contract(off)
      const int m = min(kWave, j.n - b);
      for (int k = 0; k < m; ++k) s += w[k] * (double)c[k];  // no fused multiply-add
    }
    __syncthreads();
  }
  if (t == 0) out[blockIdx.x] = s;
}

// Root-sum staging: up to kRootSumsMax descriptors, then as many results
// (the engine's pinned host staging, and the fixed device array of the
// no-allocation call, have this layout).
constexpr int kRootSumsMax = 4096;
constexpr size_t kSumJobBytes = sizeof(RootSumJob) * kRootSumsMax;
constexpr size_t kSumStageBytes = kSumJobBytes + sizeof(double) * kRootSumsMax;
static_assert(kSumJobBytes % 256 == 0, "results start 256-byte aligned");

}  // namespace

// Staging for ppsg_batch_root_sums_noalloc: the callers that cannot allocate
// get this fixed array (each GPU holds its own copy for the life of the
// process).
static __device__ __align__(256) char scratch[kSumStageBytes];

namespace {

// This device's address of `scratch` (the host-side name is not a device
// address), looked up once per device with that device current.
std::mutex g_scratchMu;
char* g_scratchAddr[64] = {};

int scratchAddress(char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail("scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratchAddr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(std::string("scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_scratchAddr[dev] = static_cast<char*>(p);
  }
  *out = g_scratchAddr[dev];
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Engine: one arena, any number of spots.
//
// Spot lifecycle: add -> live (items appended to the merged lists); park /
// remove -> alive flag cleared, items stay listed but masked; compaction
// (when masked items outnumber live ones, or when an add needs the room)
// rebuilds the lists from the live spots and frees removed spots' chunks.
// Any change re-captures the iteration graph on the next step.
// ---------------------------------------------------------------------------

namespace {

// Copies and fills in an engine's stream order, waited for on that stream.
// Nothing goes through the legacy null stream: a graph capture running in
// another thread (another engine's) makes null-stream work fail instead of
// waiting on it.
int copySync(void* dst, const void* src, size_t n, hipMemcpyKind kind, hipStream_t s) {
  if (n == 0) return 0;
  HIP_TRY(hipMemcpyAsync(dst, src, n, kind, s));
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int fillSync(void* dst, int v, size_t n, hipStream_t s) {
  if (n == 0) return 0;
  HIP_TRY(hipMemsetAsync(dst, v, n, s));
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

}  // namespace

struct ppsg_engine {
  Arena arena;
  Ctl* ctl = nullptr;
  int maxSpots = 0;
  std::vector<Spot> spots;
  // [traverser][0 forward / 1 backward per depth, 2 fold groups, 3 showdown
  // groups, 4 wide showdown groups (two lists each: [0] the entries deeper
  // than termSplit, [1] those at depth <= termSplit)]
  std::vector<DevList> lists[2][kLists];
  int groupSize = 4;  // terminal nodes per work item of spots added next
  // Terminals at depth <= termSplit run on the second stream as soon as
  // the forward pass has produced their reach rows, beside the deeper
  // forward, terminal and backward launches (0: one stream)
  int termSplit = 0;
  hipStream_t side = nullptr;
  hipEvent_t forkEv = nullptr, joinEv = nullptr;
  // per-depth launch shapes (combos per lane), [traverser][depth]: forward
  // decide, forward chance, backward decide
  std::vector<uint8_t> shapeFD[2], shapeFC[2], shapeB[2];
  int shapeOf(int tr, int dep, uint8_t kind) const {
    const std::vector<uint8_t>* v = kind == kFwdDecide   ? &shapeFD[tr]
                                    : kind == kFwdChance ? &shapeFC[tr]
                                    : kind == kBwdDecide ? &shapeB[tr]
                                                         : nullptr;
    if (!v) return 1;  // backward chance: one combo per lane
    if (dep < (int)v->size()) return (*v)[dep];
    return kind == kFwdDecide ? kRDecide : kind == kFwdChance ? kRChance : kRBackward;
  }
  void ensureShapes(int D) {
    for (int tr = 0; tr < 2; ++tr) {
      if ((int)shapeFD[tr].size() < D) shapeFD[tr].resize(D, kRDecide);
      if ((int)shapeFC[tr].size() < D) shapeFC[tr].resize(D, kRChance);
      if ((int)shapeB[tr].size() < D) shapeB[tr].resize(D, kRBackward);
    }
  }
  // device compaction scratch (grow-only)
  CompactJob* cjobs = nullptr;
  int64_t cjobsCap = 0;
  int* ccounts = nullptr;
  int64_t ccountsCap = 0;
  int* ctotals = nullptr;
  int64_t listedItems = 0, deadItems = 0;
  std::vector<int> live;
  int* liveDev = nullptr;
  int64_t liveCap = 0;
  bool graphDirty = true;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // one-spot engines (ppsg_create): kGraphIters iterations in one graph,
  // captured once the one-iteration graph has been replayed by a second
  // iterate call without a change
  bool multiIter = false;
  hipGraph_t graphK = nullptr;
  hipGraphExec_t execK = nullptr;
  int itersK = 0;  // iterations in graphK
  int stableCalls = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double lastMs = -1.0;
  int* flagsDev = nullptr;  // maxSpots alive flags staged for setAliveFlags
  int forceMask = 0;        // kernels check alive flags even with no dead items
  // root value sums: descriptors + results staged here (grow-only)
  char* sumStage = nullptr;
  int64_t sumStageCap = 0;
  char* scratchDev = nullptr;  // this device's address of `scratch` (set at init)
  char* sumHost = nullptr;     // pinned host staging of the root sums (kSumStageBytes, set at init)
  Item* walkBuf = nullptr;     // items of a subtree walk (grow-only)
  int64_t walkCap = 0;

  Dev dev() const { return Dev{arena.base, ctl}; }
  bool valid(int slot) const {
    return slot >= 0 && slot < maxSpots && spots[slot].state != kFree &&
           spots[slot].state != kDoomed;
  }
  bool isLive(int slot) const { return valid(slot) && spots[slot].state == kLive; }
  int depthCount() const { return (int)lists[0][0].size(); }
  int masked() const { return (forceMask || deadItems > 0) ? 1 : 0; }
  int init(int64_t bytes, int spotsCap);
  int add(const ppsg_tree_desc* t, int* slot);
  int remove(int slot);
  int setActive(int slot, bool on);
  int setAlive(int slot, int v);
  int setAliveAll(const std::vector<int>& flags);
  std::vector<Item> spotItems(int slot, int tr, int ps, int dep) const;
  int64_t countItems(int slot) const;
  int appendSpot(int slot);
  int compact();
  int rebuild();
  int rebuildList(int tr, int ps, int dep);
  int prepare();
  void dropGraph();
  int launchFwd(int tr, int dpt, int fromStrat, int kindMask, hipStream_t s);
  int launchTerm(int tr, int ps, int part, hipStream_t s);
  int launchBwd(int tr, int dpt, int mode, int kindMask, hipStream_t s, int dry = 0);
  int launchHalf(int tr, int mode, hipStream_t s, int dry = 0);
  int captureIters(int iters, hipGraph_t* g, hipGraphExec_t* x);
  // entries of group g (of spot S) at depth <= termSplit: a prefix of it
  int shallowCount(const Spot& S, const TermGroup& g) const {
    const auto b = S.termDepth.begin() + g.first;
    return (int)(std::upper_bound(b, b + g.count, (int16_t)std::min(termSplit, 32767)) - b);
  }
  int capture();
  int covers(int64_t iters);
  int64_t kernelsPerIter() const;
  ~ppsg_engine();
};

void ppsg_engine::dropGraph() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (execK) (void)hipGraphExecDestroy(execK);
  if (graphK) (void)hipGraphDestroy(graphK);
  exec = nullptr;
  graph = nullptr;
  execK = nullptr;
  graphK = nullptr;
  stableCalls = 0;
  graphDirty = true;
}

ppsg_engine::~ppsg_engine() {
  dropGraph();
  for (int tr = 0; tr < 2; ++tr)
    for (int ps = 0; ps < kLists; ++ps)
      for (DevList& l : lists[tr][ps]) {
        if (l.buf) (void)hipFree(l.buf);
        if (l.spare) (void)hipFree(l.spare);
      }
  for (Spot& S : spots)
    if (S.coefDev) (void)hipFree(S.coefDev);
  if (liveDev) (void)hipFree(liveDev);
  if (flagsDev) (void)hipFree(flagsDev);
  if (cjobs) (void)hipFree(cjobs);
  if (ccounts) (void)hipFree(ccounts);
  if (ctotals) (void)hipFree(ctotals);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (forkEv) (void)hipEventDestroy(forkEv);
  if (joinEv) (void)hipEventDestroy(joinEv);
  if (side) (void)hipStreamDestroy(side);
  if (stream) (void)hipStreamDestroy(stream);
  if (ctl) (void)hipFree(ctl);
  if (sumStage) (void)hipFree(sumStage);
  if (sumHost) (void)hipHostFree(sumHost);
  if (walkBuf) (void)hipFree(walkBuf);
  if (arena.base) (void)hipFree(arena.base);
}

int ppsg_engine::init(int64_t bytes, int spotsCap) {
  if (bytes <= 0 || spotsCap <= 0 || spotsCap > 65536) return fail("engine: bad capacity");
  if (scratchAddress(&scratchDev) != 0) return -1;
  {
    void* hp = nullptr;
    HIP_TRY(hipHostMalloc(&hp, kSumStageBytes, hipHostMallocDefault));
    sumHost = static_cast<char*>(hp);
  }
  HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
  bytes = roundUp(bytes);
  void* p = nullptr;
  HIP_TRY(hipMalloc(&p, (size_t)bytes));
  arena.base = static_cast<char*>(p);
  arena.cap = bytes;
  arena.holes[0] = bytes;
  HIP_TRY(hipMalloc(&p, sizeof(Ctl) * (size_t)spotsCap));
  ctl = static_cast<Ctl*>(p);
  if (fillSync(ctl, 0, sizeof(Ctl) * (size_t)spotsCap, stream) != 0) return -1;
  HIP_TRY(hipMalloc(&p, sizeof(int) * (size_t)spotsCap));
  flagsDev = static_cast<int*>(p);
  maxSpots = spotsCap;
  spots.assign(spotsCap, Spot{});
  HIP_TRY(hipEventCreate(&ev0));
  HIP_TRY(hipEventCreate(&ev1));
  HIP_TRY(hipEventCreateWithFlags(&forkEv, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&joinEv, hipEventDisableTiming));
  return 0;
}

int ppsg_engine::setAlive(int slot, int v) {
  return copySync(&ctl[slot].alive, &v, sizeof(int), hipMemcpyHostToDevice, stream);
}

int ppsg_engine::setAliveAll(const std::vector<int>& flags) {
  if (copySync(flagsDev, flags.data(), sizeof(int) * (size_t)maxSpots, hipMemcpyHostToDevice,
               stream) != 0)
    return -1;
  const unsigned g = (unsigned)std::min<int64_t>(1024, (maxSpots + kBlock - 1) / kBlock);
  setAliveFlags<<<g, kBlock, 0, stream>>>(ctl, flagsDev, maxSpots);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(stream));
  return 0;
}

// A spot's work items of one list with arena offsets: forward / backward
// items of a depth, or (ps 2 / 3 / 4) its fold / showdown / wide showdown
// groups cut at the group size.
std::vector<Item> ppsg_engine::spotItems(int slot, int tr, int ps, int dep) const {
  const Spot& S = spots[slot];
  std::vector<Item> v;
  if (ps < 2) {
    if (dep > S.maxDepth) return v;
    // planned at one combo per lane: merge each node's tiles to the depth's
    // shape (tile q at shape r covers planned tiles q*r .. q*r + r - 1)
    for (const Item& p : S.items[tr][ps][dep]) {
      const int r = shapeOf(tr, dep, p.kind);
      if (p.tile % r) continue;
      Item it = p;
      it.tile = (uint8_t)(p.tile / r);
      it.r = (uint8_t)r;
      it.a += S.off;
      it.b += S.off;
      it.c += S.off;
      it.d += S.off;
      it.f = slot;
      v.push_back(it);
    }
    return v;
  }
  // terminal lists: dep 0 = the entries deeper than termSplit, 1 = the rest
  const uint8_t kind = ps == 2 ? kBwdFold : kBwdShowdown;
  for (const TermGroup& g : S.groups[tr]) {
    if (g.kind != kind || (ps == 4) != g.wide) continue;
    const int cut = shallowCount(S, g);
    const int lo = dep == 0 ? cut : 0, hi = dep == 0 ? g.count : cut;
    for (int s0 = lo; s0 < hi; s0 += S.group) {
      Item it{};
      it.kind = g.kind;
      it.a = S.off + S.L.termNodes + (int64_t)sizeof(TermNode) * (g.first + s0);
      it.b = S.off;
      it.c = S.off + g.gather;
      it.d = S.off + g.meta;
      it.n = g.nT;
      it.m = g.nO;
      it.na = g.rows;
      it.cnt = std::min(S.group, hi - s0);
      it.f = slot;
      v.push_back(it);
    }
  }
  // largest groups first: they are the long ones
  std::stable_sort(v.begin(), v.end(), [](const Item& x, const Item& y) {
    return (int64_t)x.cnt * x.m > (int64_t)y.cnt * y.m;
  });
  return v;
}

int64_t ppsg_engine::countItems(int slot) const {
  const Spot& S = spots[slot];
  int64_t n = 0;
  for (int tr = 0; tr < 2; ++tr) {
    for (int ps = 0; ps < 2; ++ps)
      for (size_t dep = 0; dep < S.items[tr][ps].size(); ++dep)
        for (const Item& it : S.items[tr][ps][dep])
          n += it.tile % shapeOf(tr, (int)dep, it.kind) == 0;
    for (const TermGroup& g : S.groups[tr]) {
      const int cut = shallowCount(S, g);
      n += (cut + S.group - 1) / S.group + (g.count - cut + S.group - 1) / S.group;
    }
  }
  return n;
}

namespace {

int uploadTo(DevList& dl, const std::vector<Item>& v, bool append, hipStream_t s,
             int lanes = kBlock) {
  const int64_t at0 = append ? dl.count : 0;
  const int64_t need = at0 + (int64_t)v.size();
  if (need > dl.cap) {
    const int64_t cap = std::max<int64_t>(need, dl.cap + dl.cap / 2);
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, (size_t)cap * sizeof(Item)));
    if (at0 && copySync(p, dl.buf, (size_t)at0 * sizeof(Item), hipMemcpyDeviceToDevice, s) != 0)
      return -1;
    if (dl.buf) (void)hipFree(dl.buf);
    dl.buf = static_cast<Item*>(p);
    dl.cap = cap;
  }
  if (copySync(dl.buf + at0, v.data(), v.size() * sizeof(Item), hipMemcpyHostToDevice, s) != 0)
    return -1;
  dl.count = need;
  for (const Item& it : v)
    if (it.kind >= kBwdFold) dl.lds = std::max(dl.lds, terminalLdsOf(it, lanes));
  return 0;
}

}  // namespace

// Appends one spot's items to the end of every merged list.
int ppsg_engine::appendSpot(int slot) {
  Spot& S = spots[slot];
  const int D = S.maxDepth + 1;
  dropGraph();
  ensureShapes(D);
  for (int tr = 0; tr < 2; ++tr)
    for (int ps = 0; ps < kLists; ++ps) {
      std::vector<DevList>& L = lists[tr][ps];
      const int nl = ps < 2 ? D : 2;
      if ((int)L.size() < nl) L.resize(nl);
      for (int dep = 0; dep < nl; ++dep) {
        const std::vector<Item> v = spotItems(slot, tr, ps, dep);
        if (!v.empty() && uploadTo(L[dep], v, true, stream, ps == 4 ? kWideBlock : kBlock) != 0)
          return -1;
      }
    }
  S.itemCount = countItems(slot);
  S.inLists = true;
  listedItems += S.itemCount;
  return 0;
}

// Drops retired and parked spots' items from every list (on the device,
// order kept: three launches for all lists, each list compacted into its
// spare buffer, which then becomes the list) and frees removed spots.
int ppsg_engine::compact() {
  dropGraph();
  std::vector<CompactJob> jobs;
  std::vector<DevList*> owners;
  int64_t blocks = 0;
  for (int tr = 0; tr < 2; ++tr)
    for (int ps = 0; ps < kLists; ++ps)
      for (DevList& l : lists[tr][ps]) {
        if (l.count == 0) continue;
        if (l.spareCap < l.count) {
          if (l.spare) (void)hipFree(l.spare);
          l.spare = nullptr;
          l.spareCap = 0;
          void* p = nullptr;
          HIP_TRY(hipMalloc(&p, (size_t)l.cap * sizeof(Item)));
          l.spare = static_cast<Item*>(p);
          l.spareCap = l.cap;
        }
        const int64_t nb = (l.count + kCompactItems - 1) / kCompactItems;
        jobs.push_back(CompactJob{l.buf, l.spare, l.count, blocks, nb});
        owners.push_back(&l);
        blocks += nb;
      }
  if (!jobs.empty()) {
    const int64_t nj = (int64_t)jobs.size();
    if (nj > cjobsCap) {
      if (cjobs) (void)hipFree(cjobs);
      if (ctotals) (void)hipFree(ctotals);
      cjobs = nullptr;
      ctotals = nullptr;
      void* p = nullptr;
      HIP_TRY(hipMalloc(&p, (size_t)nj * sizeof(CompactJob)));
      cjobs = static_cast<CompactJob*>(p);
      HIP_TRY(hipMalloc(&p, (size_t)nj * sizeof(int)));
      ctotals = static_cast<int*>(p);
      cjobsCap = nj;
    }
    if (blocks > ccountsCap) {
      if (ccounts) (void)hipFree(ccounts);
      ccounts = nullptr;
      void* p = nullptr;
      HIP_TRY(hipMalloc(&p, (size_t)blocks * sizeof(int)));
      ccounts = static_cast<int*>(p);
      ccountsCap = blocks;
    }
    if (blocks > INT32_MAX) return fail("compaction: too many items");
    HIP_TRY(hipMemcpyAsync(cjobs, jobs.data(), (size_t)nj * sizeof(CompactJob),
                           hipMemcpyHostToDevice, stream));
    compactCount<<<(unsigned)blocks, kBlock, 0, stream>>>(cjobs, (int)nj, ctl, ccounts);
    compactScan<<<(unsigned)nj, kBlock, 0, stream>>>(cjobs, ccounts, ctotals);
    compactScatter<<<(unsigned)blocks, kBlock, 0, stream>>>(cjobs, (int)nj, ctl, ccounts);
    HIP_TRY(hipGetLastError());
    std::vector<int> kept((size_t)nj, 0);
    HIP_TRY(hipMemcpyAsync(kept.data(), ctotals, (size_t)nj * sizeof(int), hipMemcpyDeviceToHost,
                           stream));
    HIP_TRY(hipStreamSynchronize(stream));
    for (size_t j = 0; j < owners.size(); ++j) {
      DevList& l = *owners[j];
      std::swap(l.buf, l.spare);
      std::swap(l.cap, l.spareCap);
      l.count = kept[j];
    }
  }
  listedItems = 0;
  deadItems = 0;
  for (int s = 0; s < maxSpots; ++s) {
    Spot& S = spots[s];
    if (S.state == kDoomed) {
      arena.release(S.off, S.bytes);
      if (S.coefDev) (void)hipFree(S.coefDev);
      S = Spot{};
    } else if (S.state == kParked) {
      S.inLists = false;
    } else if (S.state == kLive) {
      listedItems += S.itemCount;
    }
  }
  return 0;
}

// Rebuilds every list from the live spots' items (after a launch-shape
// change) and frees removed spots.
// One merged list from the live spots' items at the current shapes (the
// lists hold no dead items when this runs).
int ppsg_engine::rebuildList(int tr, int ps, int dep) {
  dropGraph();
  std::vector<Item> v;
  for (int s = 0; s < maxSpots; ++s) {
    if (spots[s].state != kLive) continue;
    const std::vector<Item> w = spotItems(s, tr, ps, dep);
    v.insert(v.end(), w.begin(), w.end());
  }
  DevList& l = lists[tr][ps][dep];
  l.lds = 0;
  return uploadTo(l, v, false, stream, kBlock);
}

int ppsg_engine::rebuild() {
  dropGraph();
  listedItems = 0;
  deadItems = 0;
  int D = 0;
  bool any = false;
  for (int s = 0; s < maxSpots; ++s) {
    Spot& S = spots[s];
    if (S.state == kDoomed) {
      arena.release(S.off, S.bytes);
      if (S.coefDev) (void)hipFree(S.coefDev);
      S = Spot{};
      continue;
    }
    S.inLists = false;
    if (S.state == kLive) {
      D = std::max(D, S.maxDepth + 1);
      any = true;
    }
  }
  ensureShapes(D);
  for (int tr = 0; tr < 2; ++tr)
    for (int ps = 0; ps < kLists; ++ps) {
      std::vector<DevList>& L = lists[tr][ps];
      const int nl = ps < 2 ? D : (any ? 2 : 0);
      for (int dep = nl; dep < (int)L.size(); ++dep) {
        if (L[dep].buf) (void)hipFree(L[dep].buf);
        if (L[dep].spare) (void)hipFree(L[dep].spare);
      }
      L.resize(nl);
      for (int dep = 0; dep < nl; ++dep) {
        std::vector<Item> v;
        for (int s = 0; s < maxSpots; ++s) {
          if (spots[s].state != kLive) continue;
          const std::vector<Item> w = spotItems(s, tr, ps, dep);
          v.insert(v.end(), w.begin(), w.end());
        }
        if (ps >= 2)  // largest groups of every spot first
          std::stable_sort(v.begin(), v.end(), [](const Item& x, const Item& y) {
            return (int64_t)x.cnt * x.m > (int64_t)y.cnt * y.m;
          });
        L[dep].lds = 0;
        if (uploadTo(L[dep], v, false, stream, ps == 4 ? kWideBlock : kBlock) != 0) return -1;
      }
    }
  for (int s = 0; s < maxSpots; ++s)
    if (spots[s].state == kLive) {
      spots[s].inLists = true;
      spots[s].itemCount = countItems(s);
      listedItems += spots[s].itemCount;
    }
  return 0;
}

int ppsg_engine::add(const ppsg_tree_desc* t, int* slot) {
  Spot S;
  SpotTables T;
  if (planSpot(t, S, T) != 0) return -1;
  int free = -1;
  for (int s = 0; s < maxSpots && free < 0; ++s)
    if (spots[s].state == kFree) free = s;
  int64_t off = arena.alloc(S.bytes);
  bool doomed = false;
  for (const Spot& x : spots) doomed |= x.state == kDoomed;
  if ((free < 0 || off < 0) && doomed) {  // removed spots still hold room
    if (off >= 0) arena.release(off, S.bytes);
    if (compact() != 0) return -1;
    for (int s = 0; s < maxSpots && free < 0; ++s)
      if (spots[s].state == kFree) free = s;
    off = arena.alloc(S.bytes);
  }
  if (free < 0) {
    if (off >= 0) arena.release(off, S.bytes);
    g_err = "batch full: no free spot slot";
    return -2;
  }
  if (off < 0) {
    g_err = "arena full: spot needs " + std::to_string(S.bytes) + " bytes, largest free block " +
            std::to_string(arena.largestHole());
    return -2;
  }
  S.off = off;
  S.group = groupSize;
  HIP_TRY(hipStreamSynchronize(stream));
  char* base = arena.base + off;
  auto put = [&](int64_t rel, const void* src, int64_t bytes) -> int {
    if (bytes > 0 && copySync(base + rel, src, (size_t)bytes, hipMemcpyHostToDevice, stream) != 0)
      return -1;
    return 0;
  };
  const Layout& L = S.L;
  int rc = 0;
  for (int p = 0; p < 2 && rc == 0; ++p) {
    if (put(L.toParent[p], t->to_parent[p], 2 * t->list_total[p]) ||
        put(L.bmap[p], t->deal_map[p], 2 * t->deal_map_total[p]) ||
        put(L.gather[p], T.gather[p].data(), (int64_t)T.gather[p].size()) ||
        put(L.cardSame[p], T.cardSame[p].data(), 4 * t->list_total[p]) ||
        put(L.sdMeta[p], t->sd_meta[p], 16 * t->list_total[p]) ||
        put(L.reach[p] + 4 * t->reach_off[p][0], t->root_reach[p], 4 * (int64_t)t->n_base[1 - p]) ||
        put(L.weight[p], t->weight[p], 8 * (int64_t)t->n_base[p]))
      rc = -1;
  }
  if (rc == 0 && put(L.dealSib, t->deal_sib, t->deal_total)) rc = -1;
  if (rc == 0 &&
      put(L.termNodes, T.termNodes.data(), (int64_t)(T.termNodes.size() * sizeof(TermNode))))
    rc = -1;
  // untrained rows: zero regret and strategy sums (uniform average strategy)
  if (rc == 0 && (fillSync(base + L.regret, 0, 4 * (size_t)t->reg_total, stream) != 0 ||
                  fillSync(base + L.strat, 0, 4 * (size_t)t->reg_total, stream) != 0))
    rc = fail("row init failed");
  Ctl c{};
  c.alive = 1;
  if (rc == 0 && copySync(ctl + free, &c, sizeof(Ctl), hipMemcpyHostToDevice, stream) != 0)
    rc = fail("control block init failed");
  if (rc != 0) {
    arena.release(off, S.bytes);
    return -1;
  }
  S.state = kLive;
  spots[free] = std::move(S);
  if (appendSpot(free) != 0) {
    // possibly listed in part: mask it and let the next compaction free it
    Spot& F = spots[free];
    F.state = kDoomed;
    if (!F.inLists) {
      F.inLists = true;
      listedItems += F.itemCount;
    }
    deadItems += F.itemCount;
    (void)setAlive(free, 0);
    return -1;
  }
  *slot = free;
  return 0;
}

int ppsg_engine::setActive(int slot, bool on) {
  if (!valid(slot)) return fail("no such spot");
  Spot& S = spots[slot];
  if ((S.state == kLive) == on) return 0;
  dropGraph();
  if (on) {
    S.state = kLive;
    if (S.inLists) {
      if (setAlive(slot, 1) != 0) return -1;
      deadItems -= S.itemCount;
    } else {
      if (setAlive(slot, 1) != 0 || appendSpot(slot) != 0) return -1;
    }
  } else {
    S.state = kParked;
    if (setAlive(slot, 0) != 0) return -1;
    if (S.inLists) deadItems += S.itemCount;
  }
  return 0;
}

int ppsg_engine::remove(int slot) {
  if (!valid(slot)) return fail("no such spot");
  Spot& S = spots[slot];
  dropGraph();
  if (S.inLists) {
    if (S.state == kLive) {
      if (setAlive(slot, 0) != 0) return -1;
      deadItems += S.itemCount;
    }
    S.state = kDoomed;  // chunk and slot held until the lists drop its items
    return 0;
  }
  arena.release(S.off, S.bytes);
  if (S.coefDev) (void)hipFree(S.coefDev);
  S = Spot{};
  return 0;
}

// Before a launch sequence: compact when masked work dominates, refresh the
// live slot list.
int ppsg_engine::prepare() {
  if (deadItems > 0 && 2 * deadItems > listedItems)
    if (compact() != 0) return -1;
  if (!graphDirty) return 0;
  live.clear();
  for (int s = 0; s < maxSpots; ++s)
    if (spots[s].state == kLive) live.push_back(s);
  if ((int64_t)live.size() > liveCap) {
    if (liveDev) (void)hipFree(liveDev);
    liveDev = nullptr;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, (size_t)maxSpots * sizeof(int)));
    liveDev = static_cast<int*>(p);
    liveCap = maxSpots;
  }
  if (!live.empty())
    if (copySync(liveDev, live.data(), live.size() * sizeof(int), hipMemcpyHostToDevice, stream) != 0)
      return -1;
  return 0;
}

int64_t ppsg_engine::kernelsPerIter() const {
  int64_t k = 1;  // counter bump
  for (int tr = 0; tr < 2; ++tr)
    for (int ps = 0; ps < kLists; ++ps)
      for (const DevList& l : lists[tr][ps]) k += l.count > 0;
  return k;
}

int ppsg_engine::launchFwd(int tr, int dpt, int fromStrat, int kindMask, hipStream_t s) {
  const DevList& l = lists[tr][0][dpt];
  if (l.count == 0) return 0;
  launchFwdShape(shapeOf(tr, dpt, kFwdDecide), shapeOf(tr, dpt, kFwdChance), s, dev(), l.buf,
                 l.count, fromStrat, kindMask, masked());
  HIP_TRY(hipGetLastError());
  return 0;
}

int ppsg_engine::launchTerm(int tr, int ps, int part, hipStream_t s) {
  if ((int)lists[tr][ps].size() <= part) return 0;
  const DevList& l = lists[tr][ps][part];
  if (l.count == 0) return 0;
  if (ps == 2) foldKernel<<<(unsigned)l.count, kBlock, l.lds, s>>>(dev(), l.buf, masked());
  else if (ps == 3) showdownKernel<<<(unsigned)l.count, kBlock, l.lds, s>>>(dev(), l.buf, masked());
  else showdownWideKernel<<<(unsigned)l.count, kWideBlock, l.lds, s>>>(dev(), l.buf, masked());
  HIP_TRY(hipGetLastError());
  return 0;
}

int ppsg_engine::launchBwd(int tr, int dpt, int mode, int kindMask, hipStream_t s, int dry) {
  const DevList& l = lists[tr][1][dpt];
  if (l.count == 0) return 0;
  launchBwdShape(mode, shapeOf(tr, dpt, kBwdDecide), s, dev(), l.buf, l.count, kindMask, masked(),
                 dry);
  HIP_TRY(hipGetLastError());
  return 0;
}

// One half-step (or walk) of traverser tr. With a terminal split k, the
// terminals at depth <= k (read first by the backward launch of depth k - 1)
// fork onto the second stream after the forward launch of depth k - 1 (the
// last one their reach rows need) and join the main stream before that
// backward launch; everything else runs in order on s. dry: backward
// launches store the old regret and strategy values (timing only).
int ppsg_engine::launchHalf(int tr, int mode, hipStream_t s, int dry) {
  const int fromStrat = mode == PPSG_MODE_CFR ? 0 : 1;
  const int D = depthCount();
  const int k = std::min(termSplit, D);
  bool split = false;
  for (int ps = 2; ps < kLists && k > 0; ++ps)
    split |= (int)lists[tr][ps].size() > 1 && lists[tr][ps][1].count > 0;
  for (int dpt = 0; dpt < D; ++dpt) {
    if (launchFwd(tr, dpt, fromStrat, -1, s) != 0) return -1;
    if (split && dpt == k - 1) {
      HIP_TRY(hipEventRecord(forkEv, s));
      HIP_TRY(hipStreamWaitEvent(side, forkEv, 0));
      for (int ps = 2; ps < kLists; ++ps)
        if (launchTerm(tr, ps, 1, side) != 0) return -1;
      HIP_TRY(hipEventRecord(joinEv, side));
    }
  }
  for (int ps = 2; ps < kLists; ++ps) {
    if (launchTerm(tr, ps, 0, s) != 0) return -1;
    if (!split && launchTerm(tr, ps, 1, s) != 0) return -1;
  }
  for (int dpt = D - 1; dpt >= 0; --dpt) {
    if (split && dpt == k - 1) HIP_TRY(hipStreamWaitEvent(s, joinEv, 0));
    if (launchBwd(tr, dpt, mode, -1, s, dry) != 0) return -1;
  }
  return 0;
}

// Captures `iters` iterations (both half-steps and the counter bump each)
// into one graph.
int ppsg_engine::captureIters(int iters, hipGraph_t* g, hipGraphExec_t* x) {
  HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int i = 0; i < iters && rc == 0; ++i) {
    rc = launchHalf(0, PPSG_MODE_CFR, stream);
    if (rc == 0) rc = launchHalf(1, PPSG_MODE_CFR, stream);
    if (rc == 0) advanceCounters<<<1, kBlock, 0, stream>>>(ctl, liveDev, (int)live.size());
  }
  hipGraph_t cg = nullptr;
  const hipError_t e = hipStreamEndCapture(stream, &cg);
  if (rc != 0) {
    if (cg) (void)hipGraphDestroy(cg);
    return rc;
  }
  if (e != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  *g = cg;
  HIP_TRY(hipGraphInstantiate(x, cg, nullptr, nullptr, 0));
  return 0;
}

int ppsg_engine::capture() {
  if (prepare() != 0) return -1;
  if (exec && !graphDirty) {
    ++stableCalls;
    return 0;
  }
  dropGraph();
  if (captureIters(1, &graph, &exec) != 0) return -1;
  graphDirty = false;
  return 0;
}

// Every CFR launch reads coef[counter] of each live spot: refuse work that
// would run past any live spot's table.
int ppsg_engine::covers(int64_t iters) {
  for (int s = 0; s < maxSpots; ++s) {
    const Spot& S = spots[s];
    if (S.state != kLive) continue;
    if (!S.coefDev) return fail("spot " + std::to_string(s) + ": no schedule set");
    if (S.counter < 0 || S.counter + iters > S.coefLen)
      return fail("spot " + std::to_string(s) + ": schedule table too short: index " +
                  std::to_string(S.counter) + " + " + std::to_string(iters) + " > " +
                  std::to_string(S.coefLen));
  }
  return 0;
}

namespace {

int setSchedule(ppsg_t* h, int slot, const float* coef, int64_t n) {
  if (!h) return fail("null handle");
  if (!coef || n <= 0) return fail("bad schedule");
  if (!h->valid(slot)) return fail("no such spot");
  Spot& S = h->spots[slot];
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (n + 1 > S.coefCap) {  // one more entry: zeros after the last index
    const int64_t cap = std::max<int64_t>(n + 1, 2 * S.coefCap);
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, (size_t)cap * 3 * sizeof(float)));
    if (S.coefDev) (void)hipFree(S.coefDev);
    S.coefDev = static_cast<float*>(p);
    S.coefCap = cap;
    // kernels read the table through the spot's control block: no re-capture
    const float* dp = S.coefDev;
    if (copySync(&h->ctl[slot].coef, &dp, sizeof(dp), hipMemcpyHostToDevice, h->stream) != 0)
      return -1;
  }
  if (copySync(S.coefDev, coef, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, h->stream) != 0 ||
      fillSync(S.coefDev + 3 * n, 0, 3 * sizeof(float), h->stream) != 0)
    return -1;
  S.coefLen = n;
  refreshCoef<<<1, 64, 0, h->stream>>>(h->ctl, slot, (long long)n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return 0;
}

int setCounter(ppsg_t* h, int slot, int64_t t) {
  if (!h) return fail("null handle");
  if (!h->valid(slot)) return fail("no such spot");
  HIP_TRY(hipStreamSynchronize(h->stream));
  const long long v = t;
  if (copySync(&h->ctl[slot].counter, &v, sizeof(v), hipMemcpyHostToDevice, h->stream) != 0)
    return -1;
  h->spots[slot].counter = t;
  refreshCoef<<<1, 64, 0, h->stream>>>(h->ctl, slot, (long long)h->spots[slot].coefLen);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return 0;
}

int getCounter(ppsg_t* h, int slot, int64_t* t) {
  if (!h || !t) return fail("null argument");
  if (!h->valid(slot)) return fail("no such spot");
  HIP_TRY(hipStreamSynchronize(h->stream));
  long long v = 0;
  if (copySync(&v, &h->ctl[slot].counter, sizeof(v), hipMemcpyDeviceToHost, h->stream) != 0)
    return -1;
  *t = v;
  return 0;
}

int resetSpot(ppsg_t* h, int slot) {
  if (!h) return fail("null handle");
  if (!h->valid(slot)) return fail("no such spot");
  const Spot& S = h->spots[slot];
  char* base = h->arena.base + S.off;
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (fillSync(base + S.L.regret, 0, 4 * (size_t)S.regTotal, h->stream) != 0 ||
      fillSync(base + S.L.strat, 0, 4 * (size_t)S.regTotal, h->stream) != 0)
    return -1;
  return setCounter(h, slot, 0);
}

int iterate(ppsg_t* h, int n) {
  if (!h) return fail("null handle");
  if (n <= 0) return 0;
  if (h->prepare() != 0) return -1;
  if (h->live.empty()) return fail("no live spot");
  if (h->covers(n) != 0) return -1;
  if (h->capture() != 0) return -1;
  // a one-spot solve replays the same lists call after call: kGraphIters
  // iterations per replay save the gaps between replays (~4 % of the plain
  // turn's iteration). Batches change their lists every few steps (spots
  // come and go), where the extra capture cost more than it saved.
  // (as many iterations as fit kGraphNodes kernels, up to kGraphIters)
  if (h->multiIter && !h->execK && h->stableCalls > 0) {
    const int k = (int)std::min<int64_t>(kGraphIters, kGraphNodes / std::max<int64_t>(h->kernelsPerIter(), 1));
    if (k >= 2 && n >= 2 * k) {
      if (h->captureIters(k, &h->graphK, &h->execK) != 0) return -1;
      h->itersK = k;
    }
  }
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  int i = 0;
  if (h->execK)
    for (; i + h->itersK <= n; i += h->itersK) HIP_TRY(hipGraphLaunch(h->execK, h->stream));
  for (; i < n; ++i) HIP_TRY(hipGraphLaunch(h->exec, h->stream));
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->lastMs = ms;
  for (int s : h->live) h->spots[s].counter += n;
  return 0;
}

int walk(ppsg_t* h, int tr, int mode, int nSlots, const int* slots, float* const* rootCfv) {
  if (!h) return fail("null handle");
  if (nSlots > 0 && (!slots || !rootCfv)) return fail("null argument");
  if ((tr != 0 && tr != 1) || (mode != PPSG_MODE_EV && mode != PPSG_MODE_BR))
    return fail("walk: bad traverser or mode");
  for (int i = 0; i < nSlots; ++i)
    if (!h->isLive(slots[i])) return fail("walk: spot " + std::to_string(slots[i]) + " is not live");
  if (h->prepare() != 0) return -1;
  if (h->live.empty()) return fail("no live spot");
  if (h->launchHalf(tr, mode, h->stream) != 0) return -1;
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (int i = 0; i < nSlots; ++i) {
    const Spot& S = h->spots[slots[i]];
    if (rootCfv[i] && copySync(rootCfv[i], h->arena.base + S.off + S.rootCfv,
                               sizeof(float) * S.n[tr], hipMemcpyDeviceToHost, h->stream) != 0)
      return -1;
  }
  return 0;
}

int walkOnly(ppsg_t* h, int tr, int mode, int nSlots, const int* slots, float* const* rootCfv);

// EV / BR walk of one spot restricted to a subtree: the forward items of
// the listed ancestors and of nodes in [lo, hi), the subtree's terminal
// nodes (each group's run of them, cut at the group size), the backward
// items of [lo, hi). Node ids are pre-order, so [lo, hi) is the subtree of lo when
// hi is one past its last descendant. Each node's arithmetic is that of the
// full walk, so the subtree's rows come out the same bit for bit; rows
// outside it are not touched except the children of the ancestors (their
// reach).
int walkRange(ppsg_t* h, int slot, int tr, int mode, int64_t lo, int64_t hi, const int64_t* path,
              int nPath) {
  if (!h) return fail("null handle");
  if (!h->valid(slot)) return fail("walk range: no spot in slot " + std::to_string(slot));
  if ((tr != 0 && tr != 1) || (mode != PPSG_MODE_EV && mode != PPSG_MODE_BR))
    return fail("walk range: traverser 0 or 1, mode EV or BR");
  const Spot& S = h->spots[slot];
  if (lo < 0 || hi > S.nodes || lo >= hi) return fail("walk range: bad node range");
  if (nPath < 0 || (nPath > 0 && !path)) return fail("null argument");
  // a large part of the tree: the resident lists (this spot's items only)
  // cost less than building and uploading a list of its size (on the large
  // flop: whole-tree walk ~11 ms per traverser; a subtree list of 40 % of
  // the nodes took ~100 ms to build and upload)
  if (8 * (hi - lo) > S.nodes) {
    float* none[1] = {nullptr};
    return walkOnly(h, tr, mode, 1, &slot, none);
  }
  auto absolute = [&](Item it) {
    it.a += S.off;
    it.b += S.off;
    it.c += S.off;
    it.d += S.off;
    it.f = slot;
    return it;
  };
  // items are in node order per depth, terminal entries per group and
  // depth, so a node range is one run of each, found by bisection
  auto run = [](const std::vector<int32_t>& ids, int64_t a, int64_t b) {
    const auto x = std::lower_bound(ids.begin(), ids.end(), (int32_t)a);
    const auto y = std::lower_bound(x, ids.end(), (int32_t)b);
    return std::make_pair((size_t)(x - ids.begin()), (size_t)(y - ids.begin()));
  };
  struct Seg {
    int64_t first, count;
    int type;  // 0 forward, 1 fold, 2 showdown, 3 wide showdown, 4 backward
    size_t lds;
  };
  std::vector<Item> all;
  std::vector<Seg> segs;
  const int D = S.maxDepth + 1;
  for (int d = 0; d < D; ++d) {
    const int64_t first = (int64_t)all.size();
    const std::vector<Item>& v = S.items[tr][0][d];
    const std::vector<int32_t>& ids = S.itemNode[tr][0][d];
    for (int k = 0; k < nPath; ++k) {  // ancestors above the range come first in node order
      if (path[k] >= lo) continue;
      const auto r = run(ids, path[k], path[k] + 1);
      for (size_t q = r.first; q < r.second; ++q) all.push_back(absolute(v[q]));
    }
    const auto r = run(ids, lo, hi);
    for (size_t q = r.first; q < r.second; ++q) all.push_back(absolute(v[q]));
    if ((int64_t)all.size() > first) segs.push_back({first, (int64_t)all.size() - first, 0, 0});
  }
  // a group's entries run by depth, then node: one bisection per depth run
  std::vector<Item> term[3];
  for (const TermGroup& g : S.groups[tr]) {
    for (int64_t a = 0; a < g.count;) {
      int64_t e = a;
      while (e < g.count && S.termDepth[g.first + e] == S.termDepth[g.first + a]) ++e;
      const std::vector<int32_t> ids(S.termNode.begin() + g.first + a,
                                     S.termNode.begin() + g.first + e);
      const auto r = run(ids, lo, hi);
      for (size_t s0 = r.first; s0 < r.second; s0 += (size_t)S.group) {
        Item it{};
        it.kind = g.kind;
        it.a = S.off + S.L.termNodes + (int64_t)sizeof(TermNode) * (g.first + a + (int64_t)s0);
        it.b = S.off;
        it.c = S.off + g.gather;
        it.d = S.off + g.meta;
        it.n = g.nT;
        it.m = g.nO;
        it.na = g.rows;
        it.cnt = (int32_t)std::min<size_t>((size_t)S.group, r.second - s0);
        it.f = slot;
        term[g.kind == kBwdFold ? 0 : g.wide ? 2 : 1].push_back(it);
      }
      a = e;
    }
  }
  for (int k = 0; k < 3; ++k) {
    if (term[k].empty()) continue;
    size_t lds = 0;
    for (const Item& it : term[k]) lds = std::max(lds, terminalLdsOf(it, k == 2 ? kWideBlock : kBlock));
    segs.push_back({(int64_t)all.size(), (int64_t)term[k].size(), 1 + k, lds});
    all.insert(all.end(), term[k].begin(), term[k].end());
  }
  for (int d = D - 1; d >= 0; --d) {
    const int64_t first = (int64_t)all.size();
    const std::vector<Item>& v = S.items[tr][1][d];
    const auto r = run(S.itemNode[tr][1][d], lo, hi);
    for (size_t q = r.first; q < r.second; ++q) all.push_back(absolute(v[q]));
    if ((int64_t)all.size() > first) segs.push_back({first, (int64_t)all.size() - first, 4, 0});
  }
  if (all.empty()) return 0;
  if ((int64_t)all.size() > h->walkCap) {
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->walkBuf) (void)hipFree(h->walkBuf);
    h->walkBuf = nullptr;
    h->walkCap = 0;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, all.size() * sizeof(Item)));
    h->walkBuf = static_cast<Item*>(p);
    h->walkCap = (int64_t)all.size();
  }
  HIP_TRY(hipMemcpyAsync(h->walkBuf, all.data(), all.size() * sizeof(Item), hipMemcpyHostToDevice,
                         h->stream));
  const Dev dv = h->dev();
  for (const Seg& sg : segs) {
    const Item* buf = h->walkBuf + sg.first;
    const int c = (int)sg.count;
    switch (sg.type) {
      case 0: launchFwdShape(1, 1, h->stream, dv, buf, sg.count, 1, -1, 0); break;
      case 1: foldKernel<<<(unsigned)c, kBlock, sg.lds, h->stream>>>(dv, buf, 0); break;
      case 2: showdownKernel<<<(unsigned)c, kBlock, sg.lds, h->stream>>>(dv, buf, 0); break;
      case 3: showdownWideKernel<<<(unsigned)c, kWideBlock, sg.lds, h->stream>>>(dv, buf, 0); break;
      default: launchBwdShape(mode, 1, h->stream, dv, buf, sg.count, -1, 0, 0);
    }
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return 0;
}

// EV / BR half-step of the listed spots only (live or parked): every other
// spot's items are masked for the walk, the live set's flags restored after.
int walkOnly(ppsg_t* h, int tr, int mode, int nSlots, const int* slots, float* const* rootCfv) {
  if (!h) return fail("null handle");
  if (nSlots <= 0 || !slots || !rootCfv) return fail("null argument");
  if ((tr != 0 && tr != 1) || (mode != PPSG_MODE_EV && mode != PPSG_MODE_BR))
    return fail("walk: bad traverser or mode");
  for (int i = 0; i < nSlots; ++i)
    if (!h->valid(slots[i])) return fail("walk: no spot in slot " + std::to_string(slots[i]));
  if (h->prepare() != 0) return -1;  // compacts (by the live flags) before they change
  for (int i = 0; i < nSlots; ++i) {
    Spot& S = h->spots[slots[i]];
    if (S.inLists) continue;
    // a parked spot dropped by a compaction: list its items again, masked
    if (h->appendSpot(slots[i]) != 0) {
      S.inLists = true;  // possibly listed in part: keep the masking on
      h->listedItems += S.itemCount;
      h->deadItems += S.itemCount;
      return -1;
    }
    h->deadItems += S.itemCount;
  }
  std::vector<int> flags((size_t)h->maxSpots, 0);
  for (int i = 0; i < nSlots; ++i) flags[slots[i]] = 1;
  int rc = h->setAliveAll(flags);
  if (rc == 0) {
    h->forceMask = 1;
    rc = h->launchHalf(tr, mode, h->stream);
    h->forceMask = 0;
  }
  for (int s = 0; s < h->maxSpots; ++s) flags[s] = h->spots[s].state == kLive ? 1 : 0;
  const int rc2 = h->setAliveAll(flags);
  if (rc != 0) return rc;
  if (rc2 != 0) return rc2;
  for (int i = 0; i < nSlots; ++i) {
    const Spot& S = h->spots[slots[i]];
    if (rootCfv[i] && copySync(rootCfv[i], h->arena.base + S.off + S.rootCfv,
                               sizeof(float) * S.n[tr], hipMemcpyDeviceToHost, h->stream) != 0)
      return -1;
  }
  return 0;
}

int64_t bufOf(const Spot& S, int buf, int64_t& total) {
  switch (buf) {
    case PPSG_REGRET: total = S.regTotal; return S.L.regret;
    case PPSG_STRAT: total = S.regTotal; return S.L.strat;
    case PPSG_REACH0: total = S.reachTotal[0]; return S.L.reach[0];
    case PPSG_REACH1: total = S.reachTotal[1]; return S.L.reach[1];
    case PPSG_CFV: total = S.cfvTotal; return S.L.cfv;
  }
  total = 0;
  return -1;
}

int transfer(ppsg_t* h, int slot, int buf, int64_t off, int64_t n, void* host, bool toDevice) {
  if (!h || (n > 0 && !host)) return fail("null argument");
  if (!h->valid(slot)) return fail("no such spot");
  const Spot& S = h->spots[slot];
  int64_t total = 0;
  const int64_t rel = bufOf(S, buf, total);
  if (rel < 0 || off < 0 || n < 0 || off + n > total) return fail("row transfer out of range");
  HIP_TRY(hipStreamSynchronize(h->stream));
  char* dev = h->arena.base + S.off + rel + 4 * off;
  if (n) {
    if (toDevice) return copySync(dev, host, (size_t)n * 4, hipMemcpyHostToDevice, h->stream);
    return copySync(host, dev, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream);
  }
  return 0;
}

// Root value sums of the listed spots, descriptors and results staged in
// `stage` (device memory: the descriptors, padded to 256 bytes, then one
// double per listed spot). Up to kRootSumsMax spots go through the engine's
// pinned host staging (asynchronous copies); more through plain host memory.
int rootSums(ppsg_t* h, int tr, int nSlots, const int* slots, double* sums, char* stage) {
  const size_t jobBytes = roundUp((int64_t)(sizeof(RootSumJob) * (size_t)nSlots));
  const bool pinned = nSlots <= kRootSumsMax;
  std::vector<RootSumJob> own;
  if (!pinned) own.resize((size_t)nSlots);
  RootSumJob* jobs = pinned ? reinterpret_cast<RootSumJob*>(h->sumHost) : own.data();
  double* res = pinned ? reinterpret_cast<double*>(h->sumHost + kSumJobBytes) : sums;
  for (int k = 0; k < nSlots; ++k) {
    const Spot& S = h->spots[slots[k]];
    jobs[k].cfv = S.off + S.rootCfv;
    jobs[k].w = S.off + S.L.weight[tr];
    jobs[k].n = S.n[tr];
    jobs[k].pad = 0;
  }
  HIP_TRY(hipMemcpyAsync(stage, jobs, sizeof(RootSumJob) * (size_t)nSlots,
                         hipMemcpyHostToDevice, h->stream));
  rootSumKernel<<<(unsigned)nSlots, kWave, 0, h->stream>>>(
      h->arena.base, reinterpret_cast<const RootSumJob*>(stage),
      reinterpret_cast<double*>(stage + jobBytes));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(res, stage + jobBytes, sizeof(double) * (size_t)nSlots,
                         hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (pinned) std::memcpy(sums, res, sizeof(double) * (size_t)nSlots);
  return 0;
}

int rootSumsArgs(ppsg_t* h, int tr, int nSlots, const int* slots, double* sums) {
  if (!h) return fail("null handle");
  if (tr != 0 && tr != 1) return fail("root sums: traverser must be 0 or 1");
  if (nSlots < 0 || (nSlots > 0 && (!slots || !sums))) return fail("null argument");
  for (int k = 0; k < nSlots; ++k)
    if (!h->valid(slots[k])) return fail("root sums: no spot in slot " + std::to_string(slots[k]));
  return 0;
}

}  // namespace

extern "C" {

const char* ppsg_last_error(void) { return g_err.c_str(); }

int ppsg_copy_probe(int64_t bytes, int reps, double* ms_per_launch) {
  if (bytes <= 0 || bytes % 16 != 0 || reps <= 0) return fail("copy probe: bad size");
  void *a = nullptr, *b = nullptr;
  HIP_TRY(hipMalloc(&a, (size_t)bytes));
  if (hipMalloc(&b, (size_t)bytes) != hipSuccess) {
    (void)hipFree(a);
    return fail("copy probe: alloc failed");
  }
  hipEvent_t e0, e1;
  int rc = 0;
  if (hipMemset(a, 1, (size_t)bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = fail("copy probe: setup failed");
  } else {
    const int64_t n = bytes / 16;
    (void)hipEventRecord(e0, nullptr);
    const unsigned g = (unsigned)((n + kBlock - 1) / kBlock);
    for (int r = 0; r < reps; ++r) copyProbe<<<g, kBlock>>>((const float4*)a, (float4*)b, n);
    (void)hipEventRecord(e1, nullptr);
    if (hipEventSynchronize(e1) != hipSuccess) rc = fail("copy probe: launch failed");
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_launch) *ms_per_launch = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

// ---- one spot: a batch of one, in slot 0 ----

// Terminal group size for an engine holding one spot, from the mean number
// of nodes per board class in its largest terminal launch (a traverser's
// fold, showdown or wide-showdown groups): more nodes per work item spread
// a class's table loads over more nodes, fewer keep the launch wide. Per
// iteration, sizes 1-64 measured on one spot each: 17 nodes per class
// (plain turn) best at 1-2, 50 (batched-turn spots) at 4, 181 (large turn)
// at 3-6, 553 (large flop) at 32-64.
static int singleSpotGroupSize(const Spot& S) {
  int64_t nodes = 0, groups = 0;
  for (int tr = 0; tr < 2; ++tr)
    for (int kind = 0; kind < 3; ++kind) {  // fold, showdown, wide showdown
      int64_t n = 0, g = 0;
      for (const TermGroup& G : S.groups[tr]) {
        const int k = G.kind == kBwdFold ? 0 : G.wide ? 2 : 1;
        if (k != kind) continue;
        n += G.count;
        ++g;
      }
      if (n > nodes) {
        nodes = n;
        groups = g;
      }
    }
  if (groups == 0) return 4;
  const int64_t perClass = nodes / groups;
  return perClass >= 400 ? 32 : perClass < 24 ? 2 : 4;
}

int ppsg_create(const ppsg_tree_desc* t, ppsg_t** out) {
  if (!t || !out) return fail("null argument");
  *out = nullptr;
  int64_t need = 0;
  if (ppsg_footprint(t, &need) != 0) return -1;
  auto* E = new ppsg_engine();
  E->multiIter = true;
  int slot = -1;
  if (E->init(need, 1) != 0 || E->add(t, &slot) != 0) {
    delete E;
    return -1;
  }
  if (ppsg_tune(E, PPSG_TUNE_GROUP, singleSpotGroupSize(E->spots[slot])) != 0) {
    delete E;
    return -1;
  }
  *out = E;
  return 0;
}

void ppsg_destroy(ppsg_t* h) { delete h; }

int ppsg_set_schedule(ppsg_t* h, const float* coef, int64_t n) { return setSchedule(h, 0, coef, n); }
int ppsg_set_counter(ppsg_t* h, int64_t t) { return setCounter(h, 0, t); }
int ppsg_get_counter(ppsg_t* h, int64_t* t) { return getCounter(h, 0, t); }
int ppsg_reset(ppsg_t* h) { return resetSpot(h, 0); }
int ppsg_iterate(ppsg_t* h, int n) { return iterate(h, n); }

int ppsg_half_step(ppsg_t* h, int tr) {
  if (!h) return fail("null handle");
  if (tr != 0 && tr != 1) return fail("half step: bad traverser");
  if (h->prepare() != 0) return -1;
  if (h->live.empty()) return fail("no live spot");
  if (h->covers(1) != 0) return -1;
  if (h->launchHalf(tr, PPSG_MODE_CFR, h->stream) != 0) return -1;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return 0;
}

int ppsg_walk(ppsg_t* h, int tr, int mode, float* root_cfv) {
  const int slot = 0;
  float* outs[1] = {root_cfv};
  return walk(h, tr, mode, 1, &slot, outs);
}

int ppsg_read(ppsg_t* h, int buf, int64_t off, int64_t n, float* host) {
  return transfer(h, 0, buf, off, n, host, false);
}

int ppsg_write(ppsg_t* h, int buf, int64_t off, int64_t n, const float* host) {
  return transfer(h, 0, buf, off, n, const_cast<float*>(host), true);
}

int ppsg_last_iterate_ms(ppsg_t* h, double* ms) {
  if (!h || !ms) return fail("null argument");
  *ms = h->lastMs;
  return 0;
}

int ppsg_info(ppsg_t* h, int64_t* out, int cap) {
  if (!h || !out) return fail("null argument");
  if (h->prepare() != 0) return -1;
  double fb = 0.0, bb = 0.0;
  for (int s : h->live) {
    fb += h->spots[s].fwdBytes;
    bb += h->spots[s].bwdBytes;
  }
  const int64_t v[6] = {h->kernelsPerIter(), h->arena.cap, h->listedItems, h->exec ? 1 : 0,
                        (int64_t)fb, (int64_t)bb};
  for (int i = 0; i < cap && i < 6; ++i) out[i] = v[i];
  return 0;
}

int ppsg_profile(ppsg_t* h, int iters, double* out, int cap) {
  if (!h || !out) return fail("null argument");
  if (iters <= 0) return fail("profile: iters must be positive");
  if (h->prepare() != 0) return -1;
  if (h->live.empty()) return fail("no live spot");
  if (h->covers(iters) != 0) return -1;
  double acc[17] = {0};
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return fail("profile: event creation failed");
  }
  auto timed = [&](auto&& launch, int slot) -> int {
    HIP_TRY(hipEventRecord(a, h->stream));
    if (launch() != 0) return -1;
    HIP_TRY(hipEventRecord(b, h->stream));
    HIP_TRY(hipEventSynchronize(b));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, a, b));
    acc[slot] += ms;
    acc[slot + 1] += 1;
    return 0;
  };
  // cap 13: terminal launches attributed; cap 17: also decide / chance
  // re-runs of the forward and backward launches
  const bool detail = cap >= 17;
  // The backward re-runs apply the regret and strategy update a second
  // time, so in detail mode every live spot's regret and strategy tables
  // and its schedule counter are saved first and put back afterwards: the
  // profiled iterations leave no trace in the rows.
  char* snap = nullptr;
  std::vector<int64_t> savedCounter;
  if (detail) {
    size_t need = 0;
    for (int s : h->live) need += 8 * (size_t)h->spots[s].regTotal;
    if (hipMalloc(&snap, std::max<size_t>(need, 256)) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      return fail("profile: no device memory to save the rows of a detailed profile");
    }
    size_t at = 0;
    for (int s : h->live) {
      const Spot& S = h->spots[s];
      const size_t n = 4 * (size_t)S.regTotal;
      HIP_TRY(hipMemcpyAsync(snap + at, h->arena.base + S.off + S.L.regret, n,
                             hipMemcpyDeviceToDevice, h->stream));
      HIP_TRY(hipMemcpyAsync(snap + at + n, h->arena.base + S.off + S.L.strat, n,
                             hipMemcpyDeviceToDevice, h->stream));
      at += 2 * n;
      savedCounter.push_back(S.counter);
    }
  }
  const int D = h->depthCount();
  int rc = 0;
  for (int it = 0; it < iters && rc == 0; ++it) {
    for (int tr = 0; tr < 2 && rc == 0; ++tr) {
      for (int dpt = 0; dpt < D && rc == 0; ++dpt) {
        if (!h->lists[tr][0][dpt].count) continue;
        rc = timed([&] { return h->launchFwd(tr, dpt, 0, -1, h->stream); }, 0);
        // attribution: forward decide / chance re-run alone (they rewrite the
        // same reach rows from the same inputs)
        for (int k = 0; k < 2 && rc == 0 && detail; ++k)
          rc = timed([&] { return h->launchFwd(tr, dpt, 0, 1 << (kFwdDecide + k), h->stream); },
                     13 + 2 * k);
      }
      double lastFwd = acc[0];
      for (int ps = 2; ps < kLists && rc == 0; ++ps)
        for (int part = 0; part < (int)h->lists[tr][ps].size() && rc == 0; ++part) {
          if (!h->lists[tr][ps][part].count) continue;
          rc = timed([&] { return h->launchTerm(tr, ps, part, h->stream); }, 0);
          // fold and showdown have launches of their own: always attributed
          if (rc == 0 && cap >= 13) {
            acc[ps == 2 ? 9 : 11] += acc[0] - lastFwd;
            acc[ps == 2 ? 10 : 12] += 1;
          }
          lastFwd = acc[0];
        }
      for (int dpt = D - 1; dpt >= 0 && rc == 0; --dpt) {
        if (!h->lists[tr][1][dpt].count) continue;
        rc = timed([&] { return h->launchBwd(tr, dpt, PPSG_MODE_CFR, -1, h->stream); }, 2);
        for (int k = 0; k < 2 && rc == 0 && detail; ++k)
          rc = timed(
              [&] {
                return h->launchBwd(tr, dpt, PPSG_MODE_CFR, 1 << (kBwdDecide + k), h->stream);
              },
              5 + 2 * k);
      }
    }
    if (rc == 0) {
      advanceCounters<<<1, kBlock, 0, h->stream>>>(h->ctl, h->liveDev, (int)h->live.size());
      if (hipGetLastError() != hipSuccess) rc = fail("profile: counter launch failed");
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (snap) {
    size_t at = 0;
    for (int s : h->live) {
      const Spot& S = h->spots[s];
      const size_t n = 4 * (size_t)S.regTotal;
      if (hipMemcpyAsync(h->arena.base + S.off + S.L.regret, snap + at, n,
                         hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
          hipMemcpyAsync(h->arena.base + S.off + S.L.strat, snap + at + n, n,
                         hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
        rc = rc ? rc : fail("profile: restoring the saved rows failed");
      at += 2 * n;
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess && rc == 0)
      rc = fail("profile: restoring the saved rows failed");
    (void)hipFree(snap);
    if (rc != 0) return rc;
    size_t k = 0;
    for (int s : h->live)
      if (setCounter(h, s, savedCounter[k++]) != 0) return -1;
  }
  if (rc != 0) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (!detail)
    for (int s : h->live) h->spots[s].counter += iters;
  acc[4] = iters;
  for (int i = 0; i < cap && i < 17; ++i) out[i] = acc[i];
  return 0;
}

int ppsg_profile_depths(ppsg_t* h, int iters, double* out, int cap) {
  if (!h || !out) return fail("null argument");
  if (iters <= 0) return fail("profile: iters must be positive");
  if (h->prepare() != 0) return -1;
  if (h->live.empty()) return fail("no live spot");
  if (h->covers(iters) != 0) return -1;
  const int D = h->depthCount();
  if (cap < 1 + 4 * D + 4) return fail("profile: output holds " + std::to_string(cap) +
                                       " values, needs " + std::to_string(1 + 4 * D + 4));
  std::vector<double> acc((size_t)(4 * D + 4), 0.0);
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return fail("profile: event creation failed");
  }
  auto timed = [&](auto&& launch, size_t slot) -> int {
    HIP_TRY(hipEventRecord(a, h->stream));
    if (launch() != 0) return -1;
    HIP_TRY(hipEventRecord(b, h->stream));
    HIP_TRY(hipEventSynchronize(b));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, a, b));
    acc[slot] += 1e3 * ms;
    return 0;
  };
  int rc = 0;
  for (int it = 0; it < iters && rc == 0; ++it) {
    for (int tr = 0; tr < 2 && rc == 0; ++tr) {
      for (int dpt = 0; dpt < D && rc == 0; ++dpt)
        if (h->lists[tr][0][dpt].count)
          rc = timed([&] { return h->launchFwd(tr, dpt, 0, -1, h->stream); },
                     (size_t)(4 * dpt + 2 * tr));
      for (int ps = 2; ps < kLists && rc == 0; ++ps)
        for (int part = 0; part < (int)h->lists[tr][ps].size() && rc == 0; ++part)
          if (h->lists[tr][ps][part].count)
            rc = timed([&] { return h->launchTerm(tr, ps, part, h->stream); },
                       (size_t)(4 * D + (ps == 2 ? 0 : 2) + tr));
      for (int dpt = D - 1; dpt >= 0 && rc == 0; --dpt)
        if (h->lists[tr][1][dpt].count)
          rc = timed([&] { return h->launchBwd(tr, dpt, PPSG_MODE_CFR, -1, h->stream); },
                     (size_t)(4 * dpt + 2 * tr + 1));
    }
    if (rc == 0) {
      advanceCounters<<<1, kBlock, 0, h->stream>>>(h->ctl, h->liveDev, (int)h->live.size());
      if (hipGetLastError() != hipSuccess) rc = fail("profile: counter launch failed");
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc != 0) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (int s : h->live) h->spots[s].counter += iters;
  out[0] = D;
  for (size_t i = 0; i < acc.size(); ++i) out[1 + i] = acc[i];
  return 0;
}

// ---- batches ----

int ppsg_footprint(const ppsg_tree_desc* desc, int64_t* bytes) {
  if (!desc || !bytes) return fail("null argument");
  Spot S;
  SpotTables T;
  if (planSpot(desc, S, T) != 0) return -1;
  *bytes = S.bytes;
  return 0;
}

int ppsg_batch_create(int64_t arena_bytes, int max_spots, ppsg_t** out) {
  if (!out) return fail("null argument");
  *out = nullptr;
  auto* E = new ppsg_engine();
  if (E->init(arena_bytes, max_spots) != 0) {
    delete E;
    return -1;
  }
  *out = E;
  return 0;
}

int ppsg_batch_add(ppsg_t* h, const ppsg_tree_desc* desc, int* slot) {
  if (!h || !desc || !slot) return fail("null argument");
  return h->add(desc, slot);
}

int ppsg_batch_remove(ppsg_t* h, int slot) {
  if (!h) return fail("null handle");
  return h->remove(slot);
}

int ppsg_batch_set_active(ppsg_t* h, int slot, int active) {
  if (!h) return fail("null handle");
  return h->setActive(slot, active != 0);
}

int ppsg_batch_set_schedule(ppsg_t* h, int slot, const float* coef, int64_t n) {
  return setSchedule(h, slot, coef, n);
}
int ppsg_batch_set_counter(ppsg_t* h, int slot, int64_t t) { return setCounter(h, slot, t); }
int ppsg_batch_get_counter(ppsg_t* h, int slot, int64_t* t) { return getCounter(h, slot, t); }
int ppsg_batch_reset(ppsg_t* h, int slot) { return resetSpot(h, slot); }
int ppsg_batch_step(ppsg_t* h, int n) { return iterate(h, n); }

int ppsg_batch_walk(ppsg_t* h, int tr, int mode, int n_slots, const int* slots,
                    float* const* root_cfv) {
  return walk(h, tr, mode, n_slots, slots, root_cfv);
}

int ppsg_batch_walk_only(ppsg_t* h, int tr, int mode, int n_slots, const int* slots,
                         float* const* root_cfv) {
  return walkOnly(h, tr, mode, n_slots, slots, root_cfv);
}

int ppsg_batch_walk_range(ppsg_t* h, int slot, int tr, int mode, int64_t lo, int64_t hi,
                          const int64_t* path, int n_path) {
  return walkRange(h, slot, tr, mode, lo, hi, path, n_path);
}

int ppsg_batch_root_sums(ppsg_t* h, int tr, int n_slots, const int* slots, double* sums) {
  if (rootSumsArgs(h, tr, n_slots, slots, sums) != 0) return -1;
  if (n_slots == 0) return 0;
  const int64_t need = roundUp((int64_t)(sizeof(RootSumJob) * (size_t)n_slots)) +
                       (int64_t)sizeof(double) * n_slots;
  if (need > h->sumStageCap) {
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->sumStage) (void)hipFree(h->sumStage);
    h->sumStage = nullptr;
    h->sumStageCap = 0;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, (size_t)need));
    h->sumStage = static_cast<char*>(p);
    h->sumStageCap = need;
  }
  return rootSums(h, tr, n_slots, slots, sums, h->sumStage);
}

int ppsg_batch_root_sums_noalloc(ppsg_t* h, int tr, int n_slots, const int* slots, double* sums) {
  if (rootSumsArgs(h, tr, n_slots, slots, sums) != 0) return -1;
  if (n_slots > kRootSumsMax)
    return fail("root sums (no allocation): at most " + std::to_string(kRootSumsMax) +
                " spots per call, " + std::to_string(n_slots) + " given");
  if (n_slots == 0) return 0;
  if (!h->scratchDev) return fail("root sums (no allocation): scratch address not set");
  return rootSums(h, tr, n_slots, slots, sums, h->scratchDev);
}

int ppsg_batch_read(ppsg_t* h, int slot, int buf, int64_t off, int64_t n, float* host) {
  return transfer(h, slot, buf, off, n, host, false);
}

int ppsg_batch_write(ppsg_t* h, int slot, int buf, int64_t off, int64_t n, const float* host) {
  return transfer(h, slot, buf, off, n, const_cast<float*>(host), true);
}

int ppsg_batch_free_bytes(ppsg_t* h, int64_t* total, int64_t* largest) {
  if (!h) return fail("null handle");
  if (total) *total = h->arena.freeBytes();
  if (largest) *largest = h->arena.largestHole();
  return 0;
}

int ppsg_batch_compact(ppsg_t* h, double max_dead_fraction, int* compacted) {
  if (!h) return fail("null handle");
  const bool go = h->deadItems > 0 && (double)h->deadItems > max_dead_fraction * (double)h->listedItems;
  if (go && h->compact() != 0) return -1;
  if (compacted) *compacted = go ? 1 : 0;
  return 0;
}

int ppsg_kind_bytes(ppsg_t* h, double out[6]) {
  if (!h || !out) return fail("null argument");
  if (h->prepare() != 0) return -1;
  for (int k = 0; k < 6; ++k) out[k] = 0.0;
  for (int s : h->live)
    for (int k = 0; k < 6; ++k) out[k] += h->spots[s].kindBytes[k];
  return 0;
}

int ppsg_tune(ppsg_t* h, int param, int value) {
  if (!h) return fail("null handle");
  if (param == PPSG_TUNE_GROUP || param == PPSG_TUNE_ADD_GROUP) {
    if (value < 1 || value > kMaxGroup) return fail("tune: group size must be 1..64");
    h->groupSize = value;
    if (param == PPSG_TUNE_ADD_GROUP) return 0;  // spots added from now on
    bool same = true;
    for (Spot& S : h->spots) {
      same &= S.state == kFree || S.group == value;
      S.group = value;
    }
    return same ? 0 : h->rebuild();
  }
  if (param == PPSG_TUNE_TERM_SPLIT) {
    if (value < 0 || value > 32767) return fail("tune: terminal split depth must be 0..32767");
    if (value == h->termSplit) return 0;
    h->termSplit = value;
    return h->rebuild();
  }
  return fail("tune: unknown parameter");
}

int ppsg_autotune_param(ppsg_t* h, int param, const int* candidates, int n, int reps, int* chosen,
                        double* ms) {
  if (!h || (n > 0 && !candidates)) return fail("null argument");
  if (param != PPSG_TUNE_GROUP && param != PPSG_TUNE_TERM_SPLIT)
    return fail("autotune: unknown parameter");
  if (n <= 0 || reps <= 0) return fail("autotune: nothing to try");
  if (h->prepare() != 0) return -1;
  if (h->live.empty()) return fail("no live spot");
  int best = param == PPSG_TUNE_GROUP ? h->groupSize : h->termSplit;
  double bestMs = 1e300;
  for (int i = 0; i < n; ++i) {
    if (ppsg_tune(h, param, candidates[i]) != 0) return -1;
    if (h->prepare() != 0) return -1;
    // the group size: EV walks (they read the strategy sums only); the
    // split: CFR half-steps with dry backward updates (the update's loads
    // and stores, the old values stored back), the work it overlaps. Neither
    // changes a row the solve keeps (reach and cfv rows are rewritten by
    // every half-step before they are read).
    const int mode = param == PPSG_TUNE_GROUP ? PPSG_MODE_EV : PPSG_MODE_CFR;
    const int dry = param == PPSG_TUNE_GROUP ? 0 : 1;
    if (h->launchHalf(0, mode, h->stream, dry) != 0) return -1;  // warm
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    for (int r = 0; r < reps; ++r)
      for (int tr = 0; tr < 2; ++tr)
        if (h->launchHalf(tr, mode, h->stream, dry) != 0) return -1;
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipEventSynchronize(h->ev1));
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, h->ev0, h->ev1));
    if (ms) ms[i] = t / reps;
    if (t < bestMs) {
      bestMs = t;
      best = candidates[i];
    }
  }
  if (ppsg_tune(h, param, best) != 0) return -1;
  if (chosen) *chosen = best;
  return 0;
}

int ppsg_autotune(ppsg_t* h, const int* candidates, int n, int reps, int* chosen, double* ms) {
  return ppsg_autotune_param(h, PPSG_TUNE_GROUP, candidates, n, reps, chosen, ms);
}

int ppsg_autotune_shapes(ppsg_t* h, int reps, int* table, int cap) {
  if (!h) return fail("null handle");
  if (reps <= 0) return fail("autotune: reps must be positive");
  if (h->prepare() != 0) return -1;
  if (h->live.empty()) return fail("no live spot");
  if (h->deadItems > 0 && h->compact() != 0) return -1;
  const int D = h->depthCount();
  if (table && cap < 1 + 6 * D) return fail("autotune: the table needs " + std::to_string(1 + 6 * D) + " entries");
  h->ensureShapes(D);
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return fail("autotune: event creation failed");
  }
  auto mean = [&](auto&& launch) -> double {  // one untimed launch, then reps timed
    if (launch() != 0) return -1.0;
    if (hipEventRecord(a, h->stream) != hipSuccess) return -1.0;
    for (int r = 0; r < reps; ++r)
      if (launch() != 0) return -1.0;
    if (hipEventRecord(b, h->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess) return -1.0;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.0;
    return ms / reps;
  };
  // picks the fastest candidate for one shape of one list; the forward
  // launches rewrite reach rows only, the backward ones run dry (CFR
  // traffic, no row changed), so timing leaves every spot as it was
  auto pick = [&](std::vector<uint8_t>& shape, int tr, int ps, int d, const int* cand, int nc,
                  const std::function<int()>& launch) -> int {
    double best = -1.0;
    int bestR = shape[d];
    for (int k = 0; k < nc; ++k) {
      shape[d] = (uint8_t)cand[k];
      if (h->rebuildList(tr, ps, d) != 0) return -1;
      const double t = mean(launch);
      if (t < 0.0) return fail("autotune: timed launch failed");
      if (best < 0.0 || t < best) {
        best = t;
        bestR = cand[k];
      }
    }
    shape[d] = (uint8_t)bestR;
    return h->rebuildList(tr, ps, d);
  };
  int rc = 0;
  for (int tr = 0; tr < 2 && rc == 0; ++tr)
    for (int d = 0; d < D && rc == 0; ++d) {
      bool hasDec = false, hasCh = false, hasBwd = false;
      for (int s : h->live) {
        const Spot& S = h->spots[s];
        if (d > S.maxDepth) continue;
        for (const Item& it : S.items[tr][0][d]) (it.kind == kFwdDecide ? hasDec : hasCh) = true;
        for (const Item& it : S.items[tr][1][d]) hasBwd = hasBwd || it.kind == kBwdDecide;
      }
      if (hasDec)
        rc = pick(h->shapeFD[tr], tr, 0, d, kRDecideCand, 3,
                  [&] { return h->launchFwd(tr, d, 0, 1 << kFwdDecide, h->stream); });
      if (rc == 0 && hasCh)
        rc = pick(h->shapeFC[tr], tr, 0, d, kRChanceCand, 3,
                  [&] { return h->launchFwd(tr, d, 0, 1 << kFwdChance, h->stream); });
      if (rc == 0 && hasBwd)
        rc = pick(h->shapeB[tr], tr, 1, d, kRBackwardCand, 2, [&] {
          return h->launchBwd(tr, d, PPSG_MODE_CFR, 1 << kBwdDecide, h->stream, 1);
        });
    }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc != 0) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->listedItems = 0;
  for (int s : h->live) {
    h->spots[s].itemCount = h->countItems(s);
    h->listedItems += h->spots[s].itemCount;
  }
  if (table) {
    table[0] = D;
    for (int tr = 0; tr < 2; ++tr)
      for (int d = 0; d < D; ++d) {
        int* e = table + 1 + 3 * (tr * D + d);
        e[0] = h->shapeFD[tr][d];
        e[1] = h->shapeFC[tr][d];
        e[2] = h->shapeB[tr][d];
      }
  }
  return 0;
}

int ppsg_set_shapes(ppsg_t* h, const int* table, int n) {
  if (!h || !table) return fail("null argument");
  const int D = n >= 1 ? table[0] : -1;
  if (D < 0 || D > 1024 || n < 1 + 6 * D) return fail("set shapes: malformed table");
  auto in = [](int r, const int* cand, int nc) {
    for (int k = 0; k < nc; ++k)
      if (cand[k] == r) return true;
    return false;
  };
  h->ensureShapes(D);
  for (int tr = 0; tr < 2; ++tr)
    for (int d = 0; d < D; ++d) {
      const int* e = table + 1 + 3 * (tr * D + d);
      if (!in(e[0], kRDecideCand, 3) || !in(e[1], kRChanceCand, 3) || !in(e[2], kRBackwardCand, 2))
        return fail("set shapes: a shape outside the candidates (decide 1/2/4, chance 2/4/8, "
                    "backward 1/2)");
    }
  for (int tr = 0; tr < 2; ++tr)
    for (int d = 0; d < D; ++d) {
      const int* e = table + 1 + 3 * (tr * D + d);
      h->shapeFD[tr][d] = (uint8_t)e[0];
      h->shapeFC[tr][d] = (uint8_t)e[1];
      h->shapeB[tr][d] = (uint8_t)e[2];
    }
  return h->rebuild();
}

int ppsg_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
  size_t f = 0, t = 0;
  HIP_TRY(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = (int64_t)f;
  if (total_bytes) *total_bytes = (int64_t)t;
  return 0;
}

}  // extern "C"

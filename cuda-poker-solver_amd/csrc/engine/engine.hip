// MI355X DCFR engine: gfx950 kernels + the ppsg C ABI (include/ppsg.h).
//
// One DCFR iteration = two alternating half-steps (traverser 0, then 1);
// each half-step is a forward pass over depths 0..D (opponent reach) and a
// backward pass over depths D..0 (traverser counterfactual values with the
// fused DCFR update). Every (half-step, depth, pass) is one launch over a
// flat list of work items; the whole iteration (plus the schedule-counter
// bump) is captured once in a hipGraph and replayed.
//
// Work items are 64-byte descriptors built on the host with every row offset
// already resolved, so a workgroup issues its data loads after a single
// (scalar, broadcast) descriptor load instead of walking node -> class ->
// children -> offsets. Sibling rows are contiguous (child k of a node at
// base + k * stride), so one base offset reaches every child.
//
//   forward  decide (opponent acts)  one lane per opponent combo, a tile of 256
//   forward  chance (branch, tile)   child reach = parent reach / (52-board-4)
//   backward decide (tile)           RM+ value + fused DCFR update, or the
//                                    opponent-side sum of child values
//   backward chance (tile)           sum of child values over surviving branches
//   backward fold / showdown         one workgroup per node: one blocked f64
//                                    LDS scan over gathered opponent reach
// Traverser-acting decide nodes have no forward item: their children's
// reach rows alias the node's row (cuda/gpu_cfr.cu:1485-1514).
//
// No float atomics anywhere: every sum has a fixed order, so results are
// bitwise reproducible and independent of launch shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ppsg.h"

namespace {

constexpr int kBlock = 256;  // 4 wave64s
constexpr int kTile = 256;   // combos per tile item
constexpr int kMaxList = 1326;
constexpr int kRowRegs = (kMaxList + kBlock - 1) / kBlock;  // 6

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

struct Ctl {
  long long counter;  // schedule index of the running iteration
  const float* coef;  // 3 floats per schedule index
};

enum ItemKind : uint8_t {
  kFwdDecide = 0,
  kFwdChance = 1,
  kBwdDecide = 2,
  kBwdChance = 3,
  kBwdFold = 4,
  kBwdShowdown = 5,
};

// One work item. Field meaning per kind:
//   kFwdDecide   a regret/strategy row, b node reach row, c children's reach
//                block (stride n), n opponent combos, na actions
//   kFwdChance   a child class's to-parent map, b parent reach row, c child
//                reach row, n child opponent combos, v0 = 1/(52-board-4)
//   kBwdDecide   a regret/strategy row, b node cfv row, c children's cfv
//                block, m child stride, n traverser combos, e = actor is
//                the traverser
//   kBwdChance   a branch x parent-slot map (child slot or -1), b node cfv
//                row, c children's cfv block, m child stride, n traverser
//                combos (parent), na branches
//   kBwdFold /   a node reach row, b node cfv row, c gather-index list of
//   kBwdShowdown the class (opponent), d traverser list offset, e gather
//                entries per lane (8 or 16), f per-card segment table
//                index (class * 53), n traverser combos, m opponent combos,
//                v0..v2 payoffs (fold: v0), na = row registers per lane
//                (3 when both lists fit 768 combos, else 6)
struct alignas(16) Item {
  int64_t a, b, c, d;
  int32_t n, m, e, f;
  uint8_t kind, na;
  uint16_t tile;
  float v0, v1, v2;
};
static_assert(sizeof(Item) == 64, "work item must be one 64-byte line");

struct Dev {
  float* regret;
  float* strat;
  float* reach[2];
  float* cfv;
  const uint16_t* listCards[2];
  const int16_t* listSame[2];
  const uint16_t* toParent[2];
  const int16_t* branchMap[2];
  const int32_t* segOff[2];
  const uint16_t* gather[2];
  const uint32_t* sdMeta[2];
  Ctl* ctl;
};

__device__ __forceinline__ float sigmaOf(float x, float acc, float uniform, bool plus) {
  if (plus) return acc > 0.0f ? (x > 0.0f ? x / acc : 0.0f) : uniform;
  return acc > 0.0f ? x / acc : uniform;
}

// ---------------------------------------------------------------------------
// Forward.
// ---------------------------------------------------------------------------

template <int TR>
__device__ __forceinline__ void fwdItem(const Dev& d, const Item& it, int fromStrat) {
  float* __restrict__ R = d.reach[TR];
  const int lane = it.tile * kTile + (int)threadIdx.x;
  if (lane >= it.n) return;
  if (it.kind == kFwdChance) {
    const int tp = d.toParent[1 - TR][it.a + lane];
    R[it.c + lane] = R[it.b + tp] * it.v0;
    return;
  }
  const int n = it.n, na = it.na;
  const float r = R[it.b + lane];
  const float* __restrict__ src = (fromStrat ? d.strat : d.regret) + it.a + lane;
  const bool plus = !fromStrat;
  float acc = 0.0f;
  for (int a = 0; a < na; ++a) {
    const float x = src[(int64_t)a * n];
    acc += plus ? (x > 0.0f ? x : 0.0f) : x;
  }
  const float uniform = 1.0f / (float)na;
  float* __restrict__ out = R + it.c + lane;
  for (int a = 0; a < na; ++a) out[(int64_t)a * n] = r * sigmaOf(src[(int64_t)a * n], acc, uniform, plus);
}

// ---------------------------------------------------------------------------
// Backward.
// ---------------------------------------------------------------------------

// Decide tile, one lane per traverser combo. NA > 0: action count fixed at
// compile time (values stay in registers); NA == 0: generic loop.
template <int MODE, int NA>
__device__ __forceinline__ void decideTile(const Dev& d, const Item& it) {
  const int c = it.tile * kTile + (int)threadIdx.x;
  const int n = it.n;
  if (c >= n) return;
  const int na = NA > 0 ? NA : it.na;
  const int64_t stride = it.m;
  const float* __restrict__ kid = d.cfv + it.c + c;
  float* __restrict__ out = d.cfv + it.b + c;
  if (!it.e) {  // opponent acts: values add up
    float v = 0.0f;
    for (int a = 0; a < na; ++a) v += kid[a * stride];
    *out = v;
    return;
  }
  if (MODE == PPSG_MODE_BR) {
    float best = -1e30f;
    for (int a = 0; a < na; ++a) best = fmaxf(best, kid[a * stride]);
    *out = best;
    return;
  }
  constexpr bool plus = MODE == PPSG_MODE_CFR;
  float* __restrict__ src = (plus ? d.regret : d.strat) + it.a + c;
  const float uniform = 1.0f / (float)na;
  if (NA > 0) {
    float x[NA > 0 ? NA : 1], v[NA > 0 ? NA : 1], s[NA > 0 ? NA : 1];
    float acc = 0.0f;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      x[a] = src[(int64_t)a * n];
      v[a] = kid[a * stride];
      acc += plus ? (x[a] > 0.0f ? x[a] : 0.0f) : x[a];
    }
    float val = 0.0f;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      s[a] = sigmaOf(x[a], acc, uniform, plus);
      val += s[a] * v[a];
    }
    *out = val;
    if (MODE != PPSG_MODE_CFR) return;
    const float* co = d.ctl->coef + 3 * d.ctl->counter;
    const float pos = co[0], neg = co[1], avg = co[2];
    float* __restrict__ st = d.strat + it.a + c;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      st[(int64_t)a * n] = st[(int64_t)a * n] * avg + s[a];
      src[(int64_t)a * n] = x[a] * (x[a] >= 0.0f ? pos : neg) + (v[a] - val);
    }
    return;
  }
  float acc = 0.0f;
  for (int a = 0; a < na; ++a) {
    const float x = src[(int64_t)a * n];
    acc += plus ? (x > 0.0f ? x : 0.0f) : x;
  }
  float val = 0.0f;
  for (int a = 0; a < na; ++a)
    val += sigmaOf(src[(int64_t)a * n], acc, uniform, plus) * kid[a * stride];
  *out = val;
  if (MODE != PPSG_MODE_CFR) return;
  const float* co = d.ctl->coef + 3 * d.ctl->counter;
  const float pos = co[0], neg = co[1], avg = co[2];
  float* __restrict__ st = d.strat + it.a + c;
  for (int a = 0; a < na; ++a) {
    const float x = src[(int64_t)a * n];
    const float sg = sigmaOf(x, acc, uniform, plus);
    st[(int64_t)a * n] = st[(int64_t)a * n] * avg + sg;
    src[(int64_t)a * n] = x * (x >= 0.0f ? pos : neg) + (kid[a * stride] - val);
  }
}

// Chance tile, one lane per traverser (parent) combo: sum of the child
// values over the branches the combo survives, in branch order.
template <int TR>
__device__ __forceinline__ void chanceTile(const Dev& d, const Item& it) {
  const int i = it.tile * kTile + (int)threadIdx.x;
  const int n = it.n;
  if (i >= n) return;
  const int16_t* __restrict__ map = d.branchMap[TR] + it.a + i;
  const float* __restrict__ kid = d.cfv + it.c;
  const int64_t stride = it.m;
  const int nb = it.na;
  // Branches in chunks: all map loads of a chunk, then all value loads, then
  // the adds in branch order (keeps many loads in flight; same sum order).
  constexpr int kChunk = 16;
  float v = 0.0f;
  for (int b0 = 0; b0 < nb; b0 += kChunk) {
    int slot[kChunk];
    float x[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) slot[k] = b0 + k < nb ? map[(int64_t)(b0 + k) * n] : -1;
#pragma unroll
    for (int k = 0; k < kChunk; ++k) x[k] = slot[k] >= 0 ? kid[(b0 + k) * stride + slot[k]] : 0.0f;
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      if (slot[k] >= 0) v += x[k];
  }
  d.cfv[it.b + i] = v;
}

// Fold and showdown terminals. Opponent reach is staged in LDS, then each
// lane gathers its PER consecutive entries of the class's gather list
// (showdown: slots in strength order followed by the per-card segments in
// strength order; fold: the per-card segments) and one blocked f64 scan
// turns them into prefix sums X. Per traverser combo:
//   fold:     mass = total - (its two cards' segment sums) + identical combo
//   showdown: weaker / weaker-or-equal / all masses from X at le, te, n and
//             at the compile-time cut points of its two cards' segments.
// All global reads (reach, gather indices, per-combo metadata) are issued
// before the first barrier.
template <int TR, bool SD, int PER, int ROWS>
__device__ void terminalNode(const Dev& d, const Item& it, char* lds) {
  constexpr int OPP = 1 - TR;
  __shared__ double waveTot[kBlock / 64];
  __shared__ int so[53];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nO = it.m, nT = it.n;
  float* raw = reinterpret_cast<float*>(lds);
  double* X = reinterpret_cast<double*>(lds + (((nO + 1) * 4 + 15) & ~15));

  // 1. issue every global read
  const float* __restrict__ r = d.reach[TR] + it.a;
  float rv[ROWS];
#pragma unroll
  for (int q = 0; q < ROWS; ++q) {
    const int j = tid + q * kBlock;
    rv[q] = j < nO ? r[j] : 0.0f;
  }
  uint4 gi[PER / 8];
  const uint4* gsrc = reinterpret_cast<const uint4*>(d.gather[OPP] + it.c) + tid * (PER / 8);
#pragma unroll
  for (int k = 0; k < PER / 8; ++k) gi[k] = gsrc[k];
  uint4 meta[ROWS];
  uint32_t cards[ROWS];
#pragma unroll
  for (int q = 0; q < ROWS; ++q) {
    const int i = tid + q * kBlock;
    if (SD) {
      meta[q] = i < nT ? reinterpret_cast<const uint4*>(d.sdMeta[TR] + 4 * it.d)[i]
                       : make_uint4(0, 0, 0, 0);
    } else {
      cards[q] = i < nT ? ((uint32_t)d.listCards[TR][it.d + i] |
                           ((uint32_t)(d.listSame[TR][it.d + i] + 1) << 16))
                        : 0u;
    }
  }
  if (tid < 53) so[tid] = d.segOff[OPP][it.f + tid];

  // 2. stage reach (raw[nO] = 0 is the padding sentinel)
#pragma unroll
  for (int q = 0; q < ROWS; ++q) {
    const int j = tid + q * kBlock;
    if (j < nO) raw[j] = rv[q];
  }
  if (tid == 0) raw[nO] = 0.0f;
  __syncthreads();

  // 3. gather + blocked exclusive scan (f64); the lane's entries are summed
  // once for the scan and re-read from LDS when the prefixes are written
  // (keeps the lane's running values out of registers)
  double own = 0.0;
#pragma unroll
  for (int k = 0; k < PER / 8; ++k) {
    const uint32_t w[4] = {gi[k].x, gi[k].y, gi[k].z, gi[k].w};
#pragma unroll
    for (int h = 0; h < 8; ++h) own += (double)raw[(w[h >> 1] >> ((h & 1) * 16)) & 0xffffu];
  }
  double inc = own;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) waveTot[wave] = inc;
  __syncthreads();
  double base = inc - own;
  for (int w = 0; w < wave; ++w) base += waveTot[w];
  {
    double run = base;
#pragma unroll
    for (int k = 0; k < PER / 8; ++k) {
      const uint32_t w[4] = {gi[k].x, gi[k].y, gi[k].z, gi[k].w};
#pragma unroll
      for (int h = 0; h < 8; ++h) {
        X[tid * PER + k * 8 + h] = run;
        run += (double)raw[(w[h >> 1] >> ((h & 1) * 16)) & 0xffffu];
      }
    }
    if (tid == kBlock - 1) X[kBlock * PER] = run;
  }
  __syncthreads();

  // 4. per traverser combo
  float* __restrict__ out = d.cfv + it.b;
  if (SD) {
    const double* Y = X + nO;  // per-card segment prefix (offset cancels)
    const double vWin = it.v0, vTie = it.v1, vLose = it.v2;
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      const int i = tid + q * kBlock;
      if (i >= nT) continue;
      const uint4 m = meta[q];
      const int le = m.x & 0xffff, te = m.x >> 16;
      const int a1 = m.y & 0xffff, b1 = m.y >> 16;
      const int a2 = m.z & 0xffff, b2 = m.z >> 16;
      const int c1 = m.w & 0xff, c2 = (m.w >> 8) & 0xff;
      const int same = (int)(m.w >> 16) - 1;
      const double z1 = Y[so[c1]], z2 = Y[so[c2]];
      const double rs = same >= 0 ? (double)raw[same] : 0.0;
      double win = X[le] - (Y[a1] - z1) - (Y[a2] - z2);
      const double leq = X[te] - (Y[b1] - z1) - (Y[b2] - z2) + rs;
      const double all = X[nO] - (Y[so[c1 + 1]] - z1) - (Y[so[c2 + 1]] - z2) + rs;
      if (win < 0.0) win = 0.0;
      double tie = leq - win;
      if (tie < 0.0) tie = 0.0;
      double lose = all - leq;
      if (lose < 0.0) lose = 0.0;
      out[i] = (float)(win * vWin + tie * vTie + lose * vLose);
    }
  } else {
    const double tot = 0.5 * X[2 * nO];
    const double v = it.v0;
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      const int i = tid + q * kBlock;
      if (i >= nT) continue;
      const int c1 = cards[q] & 0xff, c2 = (cards[q] >> 8) & 0xff;
      const int same = (int)(cards[q] >> 16) - 1;
      double m = tot - (X[so[c1 + 1]] - X[so[c1]]) - (X[so[c2 + 1]] - X[so[c2]]);
      if (same >= 0) m += (double)raw[same];
      if (m < 0.0) m = 0.0;
      out[i] = (float)(m * v);
    }
  }
}

// Forward launch of one depth: opponent-reach items for the next depth plus
// the terminal nodes of this depth (their values need only this depth's
// reach, so they run here and the backward pass only combines).
template <int TR>
__global__ __launch_bounds__(kBlock) void fwdKernel(Dev d, const Item* __restrict__ items,
                                                    int fromStrat, int kindMask) {
  extern __shared__ __align__(16) char lds[];
  const Item it = items[blockIdx.x];
  // kindMask: profiling attribution only (-1 = every kind in real launches)
  if (!((kindMask >> it.kind) & 1)) return;
  switch (it.kind) {
    case kBwdFold:
      if (it.na == 3) {
        if (it.e == 8) terminalNode<TR, false, 8, 3>(d, it, lds);
        else terminalNode<TR, false, 16, 3>(d, it, lds);
      } else {
        if (it.e == 8) terminalNode<TR, false, 8, kRowRegs>(d, it, lds);
        else terminalNode<TR, false, 16, kRowRegs>(d, it, lds);
      }
      break;
    case kBwdShowdown:
      if (it.na == 3) {
        if (it.e == 8) terminalNode<TR, true, 8, 3>(d, it, lds);
        else terminalNode<TR, true, 16, 3>(d, it, lds);
      } else {
        if (it.e == 8) terminalNode<TR, true, 8, kRowRegs>(d, it, lds);
        else terminalNode<TR, true, 16, kRowRegs>(d, it, lds);
      }
      break;
    default:
      fwdItem<TR>(d, it, fromStrat);
  }
}

template <int TR, int MODE>
__global__ __launch_bounds__(kBlock) void bwdKernel(Dev d, const Item* __restrict__ items,
                                                    int kindMask) {
  const Item it = items[blockIdx.x];
  if (!((kindMask >> it.kind) & 1)) return;
  switch (it.kind) {
    case kBwdDecide:
      switch (it.na) {
        case 1: decideTile<MODE, 1>(d, it); break;
        case 2: decideTile<MODE, 2>(d, it); break;
        case 3: decideTile<MODE, 3>(d, it); break;
        case 4: decideTile<MODE, 4>(d, it); break;
        case 5: decideTile<MODE, 5>(d, it); break;
        case 6: decideTile<MODE, 6>(d, it); break;
        default: decideTile<MODE, 0>(d, it); break;
      }
      break;
    default:
      chanceTile<TR>(d, it);
      break;
  }
}

// Plain streaming copy (16 B per lane per step): the calibration kernel for
// hardware byte counters.
__global__ __launch_bounds__(kBlock) void copyProbe(const float4* __restrict__ a,
                                                    float4* __restrict__ b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

__global__ void advanceCounter(Ctl* ctl) {
  if (threadIdx.x == 0) ctl->counter = ctl->counter + 1;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------

struct ppsg_engine {
  Dev dev{};
  std::vector<void*> allocs;
  int64_t bytes = 0;
  int n[2] = {0, 0};
  int maxDepth = 0;
  int64_t regTotal = 0, cfvTotal = 0, reachTotal[2] = {0, 0};
  size_t terminalLds = 0;
  // per (tr, depth): device item lists + counts, and whether the forward
  // launch carries terminal nodes (it then needs dynamic LDS)
  std::vector<Item*> fwdItems[2], bwdItems[2];
  std::vector<int> fwdCount[2], bwdCount[2];
  std::vector<char> fwdTerminal[2];
  Ctl* ctl = nullptr;
  float* coef = nullptr;
  int64_t coefCap = 0;
  int64_t coefLen = 0;  // valid schedule entries
  int64_t counter = 0;  // host mirror of the device schedule index
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double lastMs = -1.0;
  int64_t kernelsPerIter = 0, itemsPerIter = 0;
  double fwdBytes = 0.0, bwdBytes = 0.0;  // algorithmic f32 row bytes / iteration

  template <typename T>
  int upload(T*& dst, const T* src, int64_t count) {
    dst = nullptr;
    if (count <= 0) return 0;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, (size_t)count * sizeof(T)));
    allocs.push_back(p);
    bytes += count * (int64_t)sizeof(T);
    if (src) HIP_TRY(hipMemcpy(p, src, (size_t)count * sizeof(T), hipMemcpyHostToDevice));
    dst = static_cast<T*>(p);
    return 0;
  }
  template <typename T>
  int uploadC(const T*& dst, const T* src, int64_t count) {
    T* p = nullptr;
    const int rc = upload(p, src, count);
    dst = p;
    return rc;
  }

  int launchFwd(int tr, int dpt, int fromStrat, int kindMask, hipStream_t s);
  int launchBwd(int tr, int dpt, int mode, int kindMask, hipStream_t s);
  int launchHalf(int tr, int mode, hipStream_t s);
  int capture();
  ~ppsg_engine();
};

ppsg_engine::~ppsg_engine() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
  for (void* p : allocs) (void)hipFree(p);
}

int ppsg_engine::launchFwd(int tr, int dpt, int fromStrat, int kindMask, hipStream_t s) {
  const int cnt = fwdCount[tr][dpt];
  if (cnt == 0) return 0;
  const size_t lds = fwdTerminal[tr][dpt] ? terminalLds : 0;
  if (tr == 0) fwdKernel<0><<<cnt, kBlock, lds, s>>>(dev, fwdItems[0][dpt], fromStrat, kindMask);
  else fwdKernel<1><<<cnt, kBlock, lds, s>>>(dev, fwdItems[1][dpt], fromStrat, kindMask);
  HIP_TRY(hipGetLastError());
  return 0;
}

int ppsg_engine::launchBwd(int tr, int dpt, int mode, int kindMask, hipStream_t s) {
  const int cnt = bwdCount[tr][dpt];
  if (cnt == 0) return 0;
  const Item* it = bwdItems[tr][dpt];
#define PPSG_BWD(T, M) bwdKernel<T, M><<<cnt, kBlock, 0, s>>>(dev, it, kindMask)
  if (tr == 0) {
    if (mode == PPSG_MODE_CFR) PPSG_BWD(0, PPSG_MODE_CFR);
    else if (mode == PPSG_MODE_EV) PPSG_BWD(0, PPSG_MODE_EV);
    else PPSG_BWD(0, PPSG_MODE_BR);
  } else {
    if (mode == PPSG_MODE_CFR) PPSG_BWD(1, PPSG_MODE_CFR);
    else if (mode == PPSG_MODE_EV) PPSG_BWD(1, PPSG_MODE_EV);
    else PPSG_BWD(1, PPSG_MODE_BR);
  }
#undef PPSG_BWD
  HIP_TRY(hipGetLastError());
  return 0;
}

int ppsg_engine::launchHalf(int tr, int mode, hipStream_t s) {
  const int fromStrat = mode == PPSG_MODE_CFR ? 0 : 1;
  for (int dpt = 0; dpt <= maxDepth; ++dpt)
    if (launchFwd(tr, dpt, fromStrat, -1, s) != 0) return -1;
  for (int dpt = maxDepth; dpt >= 0; --dpt)
    if (launchBwd(tr, dpt, mode, -1, s) != 0) return -1;
  return 0;
}

int ppsg_engine::capture() {
  if (exec) return 0;
  HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  int rc = launchHalf(0, PPSG_MODE_CFR, stream);
  if (rc == 0) rc = launchHalf(1, PPSG_MODE_CFR, stream);
  if (rc == 0) advanceCounter<<<1, 64, 0, stream>>>(ctl);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(stream, &g);
  if (rc != 0) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  graph = g;
  HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  return 0;
}

namespace {

// Gather entries per lane for a list of `len` entries: 8 or 16 (one or two
// 16-byte index loads).
int perLane(int len) { return len <= kBlock * 8 ? 8 : 16; }

// Per-class gather lists (fold: per-card segments; showdown: strength order
// followed by the per-card segments), padded to kBlock * per entries with the
// sentinel slot n (a zero in LDS).
struct GatherTables {
  std::vector<uint16_t> data[2];
  std::map<int, int64_t> foldOff[2], sdOff[2];
  std::map<int, int> foldPer[2], sdPer[2];

  void add(const ppsg_tree_desc* t, int p, int cl, bool sd) {
    auto& offs = sd ? sdOff[p] : foldOff[p];
    if (offs.count(cl)) return;
    const int n = t->cls_n[p][cl];
    const int64_t lo = t->cls_list_off[p][cl];
    std::vector<uint16_t> list;
    if (sd) list.insert(list.end(), t->sorted[p] + lo, t->sorted[p] + lo + n);
    list.insert(list.end(), t->seg_slot[p] + 2 * lo, t->seg_slot[p] + 2 * lo + 2 * n);
    const int per = perLane((int)list.size());
    list.resize((size_t)kBlock * per, (uint16_t)n);
    while (data[p].size() % 8) data[p].push_back(0);  // 16-byte aligned lanes
    offs[cl] = (int64_t)data[p].size();
    (sd ? sdPer[p] : foldPer[p])[cl] = per;
    data[p].insert(data[p].end(), list.begin(), list.end());
  }
};

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
    for (int r = 0; r < reps; ++r) copyProbe<<<4096, kBlock>>>((const float4*)a, (float4*)b, n);
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

int ppsg_create(const ppsg_tree_desc* t, ppsg_t** out) {
  if (!t || !out) return fail("null argument");
  *out = nullptr;
  if (t->n_nodes <= 0 || t->n_base[0] <= 0 || t->n_base[1] <= 0) return fail("empty tree");
  auto* E = new ppsg_engine();
  auto bail = [&](int rc) {
    delete E;
    return rc;
  };
  Dev& d = E->dev;
  const int nc = t->n_class;
  const int64_t N = t->n_nodes;
  int maxList = 0;
  for (int p = 0; p < 2; ++p)
    for (int c = 0; c < nc; ++c) maxList = std::max(maxList, t->cls_n[p][c]);
  if (maxList > kMaxList) return bail(fail("combo list longer than 1326"));
  auto kidsOf = [&](int64_t u, int& b, int& e) {
    b = t->child_base[u];
    e = u + 1 < N ? t->child_base[u + 1] : (int)t->n_child;
  };

  // branch maps per parent class: [branch][parent slot] -> child slot or -1
  std::vector<int16_t> bmap[2];
  std::map<int, int64_t> bmapOff[2];
  GatherTables G;
  for (int64_t u = 0; u < N; ++u) {
    const int cl = t->cls[u];
    if (t->kind[u] == PPSG_CHANCE) {
      int b, e;
      kidsOf(u, b, e);
      for (int p = 0; p < 2; ++p) {
        if (bmapOff[p].count(cl)) continue;
        bmapOff[p][cl] = (int64_t)bmap[p].size();
        const int np = t->cls_n[p][cl];
        for (int k = b; k < e; ++k) {
          const int ccl = t->cls[t->child_flat[k]];
          const int16_t* fp = t->from_parent[p] + t->from_parent_off[p][ccl];
          bmap[p].insert(bmap[p].end(), fp, fp + np);
        }
      }
    } else if (t->kind[u] == PPSG_FOLD || t->kind[u] == PPSG_SHOWDOWN) {
      for (int p = 0; p < 2; ++p) G.add(t, p, cl, t->kind[u] == PPSG_SHOWDOWN);
    }
  }
  for (int p = 0; p < 2; ++p) {
    if (E->uploadC(d.listCards[p], t->list_cards[p], t->list_total[p]) ||
        E->uploadC(d.listSame[p], t->list_same[p], t->list_total[p]) ||
        E->uploadC(d.toParent[p], t->to_parent[p], t->list_total[p]) ||
        E->uploadC(d.branchMap[p], bmap[p].data(), std::max<int64_t>((int64_t)bmap[p].size(), 1)) ||
        E->uploadC(d.segOff[p], t->seg_off[p], (int64_t)nc * 53) ||
        E->uploadC(d.gather[p], G.data[p].data(), std::max<int64_t>((int64_t)G.data[p].size(), 8)) ||
        E->uploadC(d.sdMeta[p], t->sd_meta[p], 4 * t->list_total[p]))
      return bail(-1);
  }
  E->n[0] = t->n_base[0];
  E->n[1] = t->n_base[1];
  E->maxDepth = t->max_depth;
  E->regTotal = t->reg_total;
  E->cfvTotal = t->cfv_total;
  E->reachTotal[0] = t->reach_total[0];
  E->reachTotal[1] = t->reach_total[1];
  if (E->upload(d.regret, (const float*)nullptr, std::max<int64_t>(t->reg_total, 1)) ||
      E->upload(d.strat, (const float*)nullptr, std::max<int64_t>(t->reg_total, 1)) ||
      E->upload(d.cfv, (const float*)nullptr, t->cfv_total) ||
      E->upload(d.reach[0], (const float*)nullptr, t->reach_total[0]) ||
      E->upload(d.reach[1], (const float*)nullptr, t->reach_total[1]))
    return bail(-1);
  // untrained rows: zero regret and strategy sums (uniform average strategy)
  if (hipMemset(d.regret, 0, sizeof(float) * std::max<int64_t>(t->reg_total, 1)) != hipSuccess ||
      hipMemset(d.strat, 0, sizeof(float) * std::max<int64_t>(t->reg_total, 1)) != hipSuccess)
    return bail(fail("row init failed"));
  for (int tr = 0; tr < 2; ++tr)
    if (hipMemcpy(d.reach[tr] + t->reach_off[tr][0], t->root_reach[tr],
                  sizeof(float) * t->n_base[1 - tr], hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail("root reach upload failed"));
  {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(Ctl)) != hipSuccess) return bail(fail("ctl alloc failed"));
    E->allocs.push_back(p);
    E->ctl = static_cast<Ctl*>(p);
    d.ctl = E->ctl;
    Ctl zero{0, nullptr};
    if (hipMemcpy(E->ctl, &zero, sizeof(Ctl), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail("ctl init failed"));
  }

  // work items
  const int D = t->max_depth + 1;
  std::vector<std::vector<Item>> fwd[2], bwd[2];
  for (int tr = 0; tr < 2; ++tr) {
    fwd[tr].assign(D, {});
    bwd[tr].assign(D, {});
    E->fwdTerminal[tr].assign(D, 0);
  }
  auto tiles = [](int n) { return (n + kTile - 1) / kTile; };
  int maxPer = 8;
  for (int64_t u = 0; u < N; ++u) {
    const int dep = t->depth[u], cl = t->cls[u], k = t->kind[u];
    int kb, ke;
    kidsOf(u, kb, ke);
    for (int tr = 0; tr < 2; ++tr) {
      const int op = 1 - tr;
      const int nT = t->cls_n[tr][cl], nO = t->cls_n[op][cl];
      if (k == PPSG_DECIDE) {
        const int na = t->na[u];
        const int first = t->child_flat[kb];
        if (t->owner[u] == op) {
          for (int q = 0; q < tiles(nO); ++q) {
            Item it{};
            it.kind = kFwdDecide;
            it.a = t->reg_off[u];
            it.b = t->reach_off[tr][u];
            it.c = t->reach_off[tr][first];
            it.n = nO;
            it.na = (uint8_t)na;
            it.tile = (uint16_t)q;
            fwd[tr][dep].push_back(it);
          }
        }
        const int64_t stride = na > 1 ? t->cfv_off[t->child_flat[kb + 1]] - t->cfv_off[first] : 0;
        for (int q = 0; q < tiles(nT); ++q) {
          Item it{};
          it.kind = kBwdDecide;
          it.a = t->reg_off[u];
          it.b = t->cfv_off[u];
          it.c = t->cfv_off[first];
          it.m = (int32_t)stride;
          it.n = nT;
          it.na = (uint8_t)na;
          it.e = t->owner[u] == tr;
          it.tile = (uint16_t)q;
          bwd[tr][dep].push_back(it);
        }
      } else if (k == PPSG_CHANCE) {
        const int nb = ke - kb;
        const float inv = 1.0f / (float)(52 - t->cls_board_n[cl] - 4);
        for (int b = kb; b < ke; ++b) {
          const int ch = t->child_flat[b];
          const int ccl = t->cls[ch];
          for (int q = 0; q < tiles(t->cls_n[op][ccl]); ++q) {
            Item it{};
            it.kind = kFwdChance;
            it.a = t->cls_list_off[op][ccl];
            it.b = t->reach_off[tr][u];
            it.c = t->reach_off[tr][ch];
            it.n = t->cls_n[op][ccl];
            it.v0 = inv;
            it.tile = (uint16_t)q;
            fwd[tr][dep].push_back(it);
          }
        }
        const int first = t->child_flat[kb];
        const int64_t stride = nb > 1 ? t->cfv_off[t->child_flat[kb + 1]] - t->cfv_off[first] : 0;
        for (int q = 0; q < tiles(nT); ++q) {
          Item it{};
          it.kind = kBwdChance;
          it.a = bmapOff[tr][cl];
          it.b = t->cfv_off[u];
          it.c = t->cfv_off[first];
          it.m = (int32_t)stride;
          it.n = nT;
          it.na = (uint8_t)nb;
          it.tile = (uint16_t)q;
          bwd[tr][dep].push_back(it);
        }
      } else {
        const bool sd = k == PPSG_SHOWDOWN;
        Item it{};
        it.kind = sd ? kBwdShowdown : kBwdFold;
        it.a = t->reach_off[tr][u];
        it.b = t->cfv_off[u];
        it.c = sd ? G.sdOff[op][cl] : G.foldOff[op][cl];
        it.d = t->cls_list_off[tr][cl];
        it.e = sd ? G.sdPer[op][cl] : G.foldPer[op][cl];
        it.f = cl * 53;
        it.n = nT;
        it.m = nO;
        it.na = (uint8_t)(std::max(nT, nO) <= 3 * kBlock ? 3 : kRowRegs);  // lanes' row registers
        const float* tv = t->term_val + u * 8 + tr * 4;
        it.v0 = tv[0];
        it.v1 = tv[1];
        it.v2 = tv[2];
        fwd[tr][dep].push_back(it);
        E->fwdTerminal[tr][dep] = 1;
        maxPer = std::max(maxPer, it.e);
      }
    }
  }
  // Algorithmic bytes of one iteration, split by pass: every f32 row element
  // read or written once per use (index tables excluded).
  for (int64_t u = 0; u < N; ++u) {
    const int cl = t->cls[u];
    const double n0 = t->cls_n[0][cl], n1 = t->cls_n[1][cl];
    if (t->kind[u] == PPSG_DECIDE) {
      const double a = t->na[u];
      const double np = t->owner[u] == 0 ? n0 : n1, nq = t->owner[u] == 0 ? n1 : n0;
      E->fwdBytes += 4 * np * (2 * a + 1);
      E->bwdBytes += 4 * np * (5 * a + 1) + 4 * nq * (a + 1);
    } else if (t->kind[u] == PPSG_CHANCE) {
      int b, e;
      kidsOf(u, b, e);
      for (int k = b; k < e; ++k) {
        const int ccl = t->cls[t->child_flat[k]];
        const double nck = (double)t->cls_n[0][ccl] + t->cls_n[1][ccl];
        E->fwdBytes += 8 * nck;
        E->bwdBytes += 4 * nck;
      }
      E->bwdBytes += 4 * (n0 + n1);
    } else {
      E->bwdBytes += 8 * (n0 + n1);
    }
  }
  for (int tr = 0; tr < 2; ++tr) {
    E->fwdItems[tr].assign(D, nullptr);
    E->bwdItems[tr].assign(D, nullptr);
    E->fwdCount[tr].assign(D, 0);
    E->bwdCount[tr].assign(D, 0);
    for (int dep = 0; dep < D; ++dep) {
      // terminal workgroups first: they are the long ones
      std::stable_sort(fwd[tr][dep].begin(), fwd[tr][dep].end(), [](const Item& a, const Item& b) {
        return a.kind >= kBwdFold && b.kind < kBwdFold;
      });
      E->fwdCount[tr][dep] = (int)fwd[tr][dep].size();
      E->bwdCount[tr][dep] = (int)bwd[tr][dep].size();
      if (E->upload(E->fwdItems[tr][dep], fwd[tr][dep].data(), (int64_t)fwd[tr][dep].size()) ||
          E->upload(E->bwdItems[tr][dep], bwd[tr][dep].data(), (int64_t)bwd[tr][dep].size()))
        return bail(-1);
      E->kernelsPerIter += (fwd[tr][dep].empty() ? 0 : 1) + (bwd[tr][dep].empty() ? 0 : 1);
      E->itemsPerIter += (int64_t)(fwd[tr][dep].size() + bwd[tr][dep].size());
    }
  }
  E->kernelsPerIter += 1;
  // LDS of one terminal workgroup: reach (+ sentinel) and the f64 prefix array
  E->terminalLds = (size_t)(((maxList + 1) * 4 + 15) & ~15) + (size_t)(kBlock * maxPer + 1) * 8;
  if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&E->ev0) != hipSuccess || hipEventCreate(&E->ev1) != hipSuccess)
    return bail(fail("stream/event creation failed"));
  if (E->terminalLds > 64 * 1024) {
    const void* ks[] = {(const void*)fwdKernel<0>, (const void*)fwdKernel<1>};
    for (const void* k : ks)
      if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)E->terminalLds) != hipSuccess)
        return bail(fail("cannot raise dynamic LDS limit"));
  }
  *out = E;
  return 0;
}

void ppsg_destroy(ppsg_t* h) { delete h; }

int ppsg_set_schedule(ppsg_t* h, const float* coef, int64_t n) {
  if (!h || !coef || n <= 0) return fail("bad schedule");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (n > h->coefCap) {
    void* p = nullptr;
    const int64_t cap = std::max<int64_t>(n, 2 * h->coefCap);
    HIP_TRY(hipMalloc(&p, (size_t)cap * 3 * sizeof(float)));
    h->allocs.push_back(p);
    h->bytes += cap * 12;
    h->coef = static_cast<float*>(p);
    h->coefCap = cap;
    // kernels read the table through the control block: no re-capture
    HIP_TRY(hipMemcpy(&h->ctl->coef, &h->coef, sizeof(float*), hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(h->coef, coef, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice));
  h->coefLen = n;
  return 0;
}

// Every CFR launch reads coef[counter]: refuse work that would run past the table.
static int scheduleCovers(ppsg_t* h, int64_t iters) {
  if (!h->coef) return fail("no schedule set");
  if (h->counter < 0 || h->counter + iters > h->coefLen)
    return fail("schedule table too short: index " + std::to_string(h->counter) + " + " +
                std::to_string(iters) + " > " + std::to_string(h->coefLen));
  return 0;
}

int ppsg_set_counter(ppsg_t* h, int64_t t) {
  if (!h) return fail("null handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  long long v = t;
  HIP_TRY(hipMemcpy(&h->ctl->counter, &v, sizeof(v), hipMemcpyHostToDevice));
  h->counter = t;
  return 0;
}

int ppsg_get_counter(ppsg_t* h, int64_t* t) {
  if (!h || !t) return fail("null argument");
  HIP_TRY(hipStreamSynchronize(h->stream));
  long long v = 0;
  HIP_TRY(hipMemcpy(&v, &h->ctl->counter, sizeof(v), hipMemcpyDeviceToHost));
  *t = v;
  return 0;
}

int ppsg_reset(ppsg_t* h) {
  if (!h) return fail("null handle");
  HIP_TRY(hipMemsetAsync(h->dev.regret, 0, (size_t)std::max<int64_t>(h->regTotal, 1) * 4, h->stream));
  HIP_TRY(hipMemsetAsync(h->dev.strat, 0, (size_t)std::max<int64_t>(h->regTotal, 1) * 4, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return ppsg_set_counter(h, 0);
}

int ppsg_iterate(ppsg_t* h, int n) {
  if (!h) return fail("null handle");
  if (n <= 0) return 0;
  if (scheduleCovers(h, n) != 0) return -1;
  if (h->capture() != 0) return -1;
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  for (int i = 0; i < n; ++i) HIP_TRY(hipGraphLaunch(h->exec, h->stream));
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->lastMs = ms;
  h->counter += n;
  return 0;
}

int ppsg_half_step(ppsg_t* h, int tr) {
  if (!h) return fail("null handle");
  if (tr != 0 && tr != 1) return fail("half step: bad traverser");
  if (scheduleCovers(h, 1) != 0) return -1;
  if (h->launchHalf(tr, PPSG_MODE_CFR, h->stream) != 0) return -1;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return 0;
}

int ppsg_walk(ppsg_t* h, int tr, int mode, float* root_cfv) {
  if (!h) return fail("null handle");
  if ((tr != 0 && tr != 1) || (mode != PPSG_MODE_EV && mode != PPSG_MODE_BR))
    return fail("walk: bad traverser or mode");
  if (h->launchHalf(tr, mode, h->stream) != 0) return -1;
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (root_cfv)
    HIP_TRY(hipMemcpy(root_cfv, h->dev.cfv, sizeof(float) * h->n[tr], hipMemcpyDeviceToHost));
  return 0;
}

static float* bufOf(ppsg_t* h, int buf, int64_t& total) {
  switch (buf) {
    case PPSG_REGRET: total = h->regTotal; return h->dev.regret;
    case PPSG_STRAT: total = h->regTotal; return h->dev.strat;
    case PPSG_REACH0: total = h->reachTotal[0]; return h->dev.reach[0];
    case PPSG_REACH1: total = h->reachTotal[1]; return h->dev.reach[1];
    case PPSG_CFV: total = h->cfvTotal; return h->dev.cfv;
  }
  total = 0;
  return nullptr;
}

int ppsg_read(ppsg_t* h, int buf, int64_t off, int64_t n, float* host) {
  if (!h || !host) return fail("null argument");
  int64_t total = 0;
  float* p = bufOf(h, buf, total);
  if (!p || off < 0 || n < 0 || off + n > total) return fail("read out of range");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (n) HIP_TRY(hipMemcpy(host, p + off, (size_t)n * 4, hipMemcpyDeviceToHost));
  return 0;
}

int ppsg_write(ppsg_t* h, int buf, int64_t off, int64_t n, const float* host) {
  if (!h || !host) return fail("null argument");
  int64_t total = 0;
  float* p = bufOf(h, buf, total);
  if (!p || off < 0 || n < 0 || off + n > total) return fail("write out of range");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (n) HIP_TRY(hipMemcpy(p + off, host, (size_t)n * 4, hipMemcpyHostToDevice));
  return 0;
}

int ppsg_last_iterate_ms(ppsg_t* h, double* ms) {
  if (!h || !ms) return fail("null argument");
  *ms = h->lastMs;
  return 0;
}

int ppsg_info(ppsg_t* h, int64_t* out, int cap) {
  if (!h || !out) return fail("null argument");
  const int64_t v[6] = {h->kernelsPerIter, h->bytes, h->itemsPerIter, h->exec ? 1 : 0,
                        (int64_t)h->fwdBytes, (int64_t)h->bwdBytes};
  for (int i = 0; i < cap && i < 6; ++i) out[i] = v[i];
  return 0;
}

int ppsg_profile(ppsg_t* h, int iters, double* out, int cap) {
  if (!h || !out) return fail("null argument");
  if (scheduleCovers(h, iters) != 0) return -1;
  double acc[13] = {0};
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
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
  int rc = 0;
  for (int it = 0; it < iters && rc == 0; ++it) {
    for (int tr = 0; tr < 2 && rc == 0; ++tr) {
      for (int dpt = 0; dpt <= h->maxDepth && rc == 0; ++dpt) {
        if (!h->fwdCount[tr][dpt]) continue;
        rc = timed([&] { return h->launchFwd(tr, dpt, 0, -1, h->stream); }, 0);
        // attribution: terminal kinds re-run alone (they only write cfv rows)
        for (int k = 0; k < 2 && rc == 0 && cap >= 13; ++k)
          rc = timed([&] { return h->launchFwd(tr, dpt, 0, 1 << (kBwdFold + k), h->stream); },
                     9 + 2 * k);
      }
      for (int dpt = h->maxDepth; dpt >= 0 && rc == 0; --dpt) {
        if (!h->bwdCount[tr][dpt]) continue;
        rc = timed([&] { return h->launchBwd(tr, dpt, PPSG_MODE_CFR, -1, h->stream); }, 2);
        // attribution: decide and chance tiles re-run alone
        for (int k = 0; k < 2 && rc == 0 && cap >= 13; ++k)
          rc = timed([&] { return h->launchBwd(tr, dpt, PPSG_MODE_CFR, 1 << (kBwdDecide + k),
                                               h->stream); },
                     5 + 2 * k);
      }
    }
    if (rc == 0) advanceCounter<<<1, 64, 0, h->stream>>>(h->ctl);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc != 0) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->counter += iters;
  acc[4] = iters;
  for (int i = 0; i < cap && i < 13; ++i) out[i] = acc[i];
  return 0;
}

}  // extern "C"

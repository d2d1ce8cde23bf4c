// MI355X DCFR engine: gfx950 kernels + the ppsg C ABI (include/ppsg.h).
//
// One DCFR iteration = two alternating half-steps (traverser 0, then 1);
// each half-step is a forward pass over depths 0..D (opponent reach) and a
// backward pass over depths D..0 (traverser counterfactual values with the
// fused DCFR update). Every (half-step, depth) is one launch over a flat
// list of work items; the whole iteration (plus the schedule-counter bump)
// is captured once in a hipGraph and replayed.
//
// Work items (int2 {node, arg}):
//   forward  decide  (opponent acts)   arg = tile        one lane per opponent combo
//   forward  chance                    arg = branch<<16 | tile
//   backward decide                    arg = tile        one lane per traverser combo
//   backward chance                    arg = tile        gather over branches
//   backward fold / showdown           arg = 0           one workgroup per node
// Traverser-acting decide nodes have no forward item: their children's reach
// rows alias the node's row (cuda/gpu_cfr.cu:1485-1514).
//
// Every reduction has a fixed order (no float atomics): results do not
// depend on launch shape, run or batch composition.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ppsg.h"

namespace {

constexpr int kBlock = 256;  // 4 wave64s
constexpr int kTile = 256;   // combos per tile item
constexpr int kMaxBranch = 64;

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

// Kernel view of one compiled tree (passed by value).
struct Dev {
  const int32_t* clsN[2];
  const int64_t* clsOff[2];
  const int32_t* clsBoardN;
  const uint16_t* listCards[2];
  const int16_t* listSame[2];
  const uint16_t* toParent[2];
  const int64_t* fromParentOff[2];
  const int16_t* fromParent[2];
  const int32_t* segOff[2];
  const uint16_t* segSlot[2];
  const uint16_t* sorted[2];
  const uint32_t* sdMeta[2];
  const uint8_t* kind;
  const int8_t* owner;
  const uint8_t* na;
  const int32_t* cls;
  const int32_t* childBase;
  const int32_t* childFlat;
  const int64_t* regOff;
  const int64_t* cfvOff;
  const int64_t* reachOff[2];
  const float* termVal;
  float* regret;
  float* strat;
  float* reach[2];
  float* cfv;
  Ctl* ctl;
};

// ---------------------------------------------------------------------------
// Strategy from a row set: regret matching+ (CFR) or the normalized
// strategy sum (EV/BR walks). Uniform when nothing is positive.
// ---------------------------------------------------------------------------

__device__ __forceinline__ float sigmaOf(float x, float acc, float uniform, bool plus) {
  if (plus) return acc > 0.0f ? (x > 0.0f ? x / acc : 0.0f) : uniform;
  return acc > 0.0f ? x / acc : uniform;
}

// ---------------------------------------------------------------------------
// Forward: opponent reach.
// ---------------------------------------------------------------------------

template <int TR>
__global__ __launch_bounds__(kBlock) void fwdKernel(Dev d, const int2* __restrict__ items,
                                                    int fromStrat) {
  constexpr int OPP = 1 - TR;
  const int2 it = items[blockIdx.x];
  const int node = it.x;
  float* __restrict__ R = d.reach[TR];
  if (d.kind[node] == PPSG_DECIDE) {
    const int n = d.clsN[OPP][d.cls[node]];
    const int c = it.y * kTile + (int)threadIdx.x;
    if (c >= n) return;
    const int na = d.na[node];
    const float r = R[d.reachOff[TR][node] + c];
    const float* __restrict__ src = (fromStrat ? d.strat : d.regret) + d.regOff[node] + c;
    const bool plus = !fromStrat;
    float acc = 0.0f;
    for (int a = 0; a < na; ++a) {
      const float x = src[(int64_t)a * n];
      acc += plus ? (x > 0.0f ? x : 0.0f) : x;
    }
    const float uniform = 1.0f / (float)na;
    const int32_t* kids = d.childFlat + d.childBase[node];
    for (int a = 0; a < na; ++a) {
      const float s = sigmaOf(src[(int64_t)a * n], acc, uniform, plus);
      R[d.reachOff[TR][kids[a]] + c] = r * s;
    }
    return;
  }
  // chance branch: child reach = parent reach / (52 - board - 4), compacted
  const int br = it.y >> 16;
  const int j = (it.y & 0xffff) * kTile + (int)threadIdx.x;
  const int child = d.childFlat[d.childBase[node] + br];
  const int ccl = d.cls[child];
  if (j >= d.clsN[OPP][ccl]) return;
  const float inv = 1.0f / (float)(52 - d.clsBoardN[d.cls[node]] - 4);
  const int tp = d.toParent[OPP][d.clsOff[OPP][ccl] + j];
  R[d.reachOff[TR][child] + j] = R[d.reachOff[TR][node] + tp] * inv;
}

// ---------------------------------------------------------------------------
// Backward pieces.
// ---------------------------------------------------------------------------

// Decide node, one lane per traverser combo. NA > 0: action count known at
// compile time (values stay in registers); NA == 0: generic loop.
template <int TR, int MODE, int NA>
__device__ __forceinline__ void decideTile(const Dev& d, int node, int tile) {
  const int n = d.clsN[TR][d.cls[node]];
  const int c = tile * kTile + (int)threadIdx.x;
  if (c >= n) return;
  const int na = NA > 0 ? NA : d.na[node];
  const int32_t* kids = d.childFlat + d.childBase[node];
  float* __restrict__ out = d.cfv + d.cfvOff[node];
  const float* __restrict__ cf = d.cfv;
  if (d.owner[node] != TR) {  // opponent acts: values add up
    float v = 0.0f;
    for (int a = 0; a < na; ++a) v += cf[d.cfvOff[kids[a]] + c];
    out[c] = v;
    return;
  }
  if (MODE == PPSG_MODE_BR) {
    float best = -1e30f;
    for (int a = 0; a < na; ++a) best = fmaxf(best, cf[d.cfvOff[kids[a]] + c]);
    out[c] = best;
    return;
  }
  const int64_t base = d.regOff[node] + c;
  float* __restrict__ src = MODE == PPSG_MODE_CFR ? d.regret : d.strat;
  constexpr bool plus = MODE == PPSG_MODE_CFR;
  const float uniform = 1.0f / (float)na;
  if (NA > 0) {
    float x[NA > 0 ? NA : 1], v[NA > 0 ? NA : 1];
    float acc = 0.0f;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      x[a] = src[base + (int64_t)a * n];
      v[a] = cf[d.cfvOff[kids[a]] + c];
      acc += plus ? (x[a] > 0.0f ? x[a] : 0.0f) : x[a];
    }
    float s[NA > 0 ? NA : 1];
    float val = 0.0f;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      s[a] = sigmaOf(x[a], acc, uniform, plus);
      val += s[a] * v[a];
    }
    out[c] = val;
    if (MODE != PPSG_MODE_CFR) return;
    const float* co = d.ctl->coef + 3 * d.ctl->counter;
    const float pos = co[0], neg = co[1], avg = co[2];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      const int64_t e = base + (int64_t)a * n;
      d.strat[e] = d.strat[e] * avg + s[a];
      d.regret[e] = x[a] * (x[a] >= 0.0f ? pos : neg) + (v[a] - val);
    }
    return;
  }
  float acc = 0.0f;
  for (int a = 0; a < na; ++a) {
    const float x = src[base + (int64_t)a * n];
    acc += plus ? (x > 0.0f ? x : 0.0f) : x;
  }
  float val = 0.0f;
  for (int a = 0; a < na; ++a)
    val += sigmaOf(src[base + (int64_t)a * n], acc, uniform, plus) * cf[d.cfvOff[kids[a]] + c];
  out[c] = val;
  if (MODE != PPSG_MODE_CFR) return;
  const float* co = d.ctl->coef + 3 * d.ctl->counter;
  const float pos = co[0], neg = co[1], avg = co[2];
  for (int a = 0; a < na; ++a) {
    const int64_t e = base + (int64_t)a * n;
    const float x = d.regret[e];
    const float s = sigmaOf(x, acc, uniform, plus);
    d.strat[e] = d.strat[e] * avg + s;
    d.regret[e] = x * (x >= 0.0f ? pos : neg) + (cf[d.cfvOff[kids[a]] + c] - val);
  }
}

// Chance node, one lane per traverser (parent) combo: sum of the child
// values over the branches the combo survives, in branch order.
template <int TR>
__device__ __forceinline__ void chanceTile(const Dev& d, int node, int tile) {
  __shared__ int64_t brCfv[kMaxBranch];
  __shared__ int64_t brMap[kMaxBranch];
  const int nb = d.na[node];
  const int32_t* kids = d.childFlat + d.childBase[node];
  if ((int)threadIdx.x < nb) {
    const int ch = kids[threadIdx.x];
    brCfv[threadIdx.x] = d.cfvOff[ch];
    brMap[threadIdx.x] = d.fromParentOff[TR][d.cls[ch]];
  }
  __syncthreads();
  const int n = d.clsN[TR][d.cls[node]];
  const int i = tile * kTile + (int)threadIdx.x;
  if (i >= n) return;
  const int16_t* __restrict__ fp = d.fromParent[TR];
  const float* __restrict__ cf = d.cfv;
  float v = 0.0f;
  for (int b = 0; b < nb; ++b) {
    const int slot = fp[brMap[b] + i];
    if (slot >= 0) v += cf[brCfv[b] + slot];
  }
  d.cfv[d.cfvOff[node] + i] = v;
}

// Exclusive prefix sum over X[0..m) in place (f64, fixed order); X[m]
// receives the total. All lanes of the workgroup must call it.
__device__ void blockScan(double* X, int m, double* waveTot) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int per = (m + kBlock - 1) / kBlock;
  const int lo = min(tid * per, m), hi = min(lo + per, m);
  double own = 0.0;
  for (int k = lo; k < hi; ++k) own += X[k];
  double inc = own;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) waveTot[wave] = inc;
  __syncthreads();
  double run = inc - own;
  for (int w = 0; w < wave; ++w) run += waveTot[w];
  for (int k = lo; k < hi; ++k) {
    const double t = X[k];
    X[k] = run;
    run += t;
  }
  __syncthreads();
  if (tid == kBlock - 1) X[m] = run;
  __syncthreads();
}

// Fold terminal: per-card opponent sums by a scan over the per-card
// segments; mass = total - both cards' sums + the identical combo.
template <int TR>
__device__ void foldNode(const Dev& d, int node, char* lds) {
  constexpr int OPP = 1 - TR;
  __shared__ double waveTot[4];
  const int cl = d.cls[node];
  const int nO = d.clsN[OPP][cl], nT = d.clsN[TR][cl];
  float* raw = reinterpret_cast<float*>(lds);
  double* X = reinterpret_cast<double*>(lds + ((nO * 4 + 15) & ~15));
  const float* __restrict__ r = d.reach[TR] + d.reachOff[TR][node];
  for (int j = threadIdx.x; j < nO; j += kBlock) raw[j] = r[j];
  __syncthreads();
  const uint16_t* seg = d.segSlot[OPP] + 2 * d.clsOff[OPP][cl];
  for (int q = threadIdx.x; q < 2 * nO; q += kBlock) X[q] = (double)raw[seg[q]];
  __syncthreads();
  blockScan(X, 2 * nO, waveTot);
  const int32_t* so = d.segOff[OPP] + (int64_t)cl * 53;
  const double tot = 0.5 * X[2 * nO];
  const double v = (double)d.termVal[(int64_t)node * 8 + TR * 4];
  const int64_t off = d.clsOff[TR][cl];
  float* out = d.cfv + d.cfvOff[node];
  for (int i = threadIdx.x; i < nT; i += kBlock) {
    const int cc = d.listCards[TR][off + i];
    const int c1 = cc & 0xff, c2 = cc >> 8;
    const int same = d.listSame[TR][off + i];
    double m = tot - (X[so[c1 + 1]] - X[so[c1]]) - (X[so[c2 + 1]] - X[so[c2]]);
    if (same >= 0) m += (double)raw[same];
    if (m < 0.0) m = 0.0;
    out[i] = (float)(m * v);
  }
}

// Showdown terminal. Opponent reach is gathered in strength order and
// scanned (S), and the per-card segments — strength-ordered — are scanned
// too (Y). For each traverser combo the compile-time bounds (le, te and
// the segment cut points) turn the weaker / weaker-or-equal / all masses
// into prefix differences with the shared-card entries removed.
template <int TR>
__device__ void showdownNode(const Dev& d, int node, char* lds) {
  constexpr int OPP = 1 - TR;
  __shared__ double waveTot[4];
  __shared__ int so[53];
  const int cl = d.cls[node];
  const int nO = d.clsN[OPP][cl], nT = d.clsN[TR][cl];
  float* raw = reinterpret_cast<float*>(lds);
  double* X = reinterpret_cast<double*>(lds + ((nO * 4 + 15) & ~15));
  const float* __restrict__ r = d.reach[TR] + d.reachOff[TR][node];
  for (int j = threadIdx.x; j < nO; j += kBlock) raw[j] = r[j];
  if (threadIdx.x < 53) so[threadIdx.x] = d.segOff[OPP][(int64_t)cl * 53 + threadIdx.x];
  __syncthreads();
  const int64_t offO = d.clsOff[OPP][cl];
  const uint16_t* srt = d.sorted[OPP] + offO;
  const uint16_t* seg = d.segSlot[OPP] + 2 * offO;
  for (int k = threadIdx.x; k < nO; k += kBlock) X[k] = (double)raw[srt[k]];
  for (int q = threadIdx.x; q < 2 * nO; q += kBlock) X[nO + q] = (double)raw[seg[q]];
  __syncthreads();
  blockScan(X, 3 * nO, waveTot);
  // X[k] = sum of the k weakest; Y(q) = X[nO + q] (offset cancels in differences)
  const double* Y = X + nO;
  const float* tv = d.termVal + (int64_t)node * 8 + TR * 4;
  const double vWin = tv[0], vTie = tv[1], vLose = tv[2];
  const int64_t off = d.clsOff[TR][cl];
  const uint4* meta = reinterpret_cast<const uint4*>(d.sdMeta[TR] + 4 * off);
  float* out = d.cfv + d.cfvOff[node];
  for (int i = threadIdx.x; i < nT; i += kBlock) {
    const uint4 m = meta[i];
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
}

template <int TR, int MODE>
__global__ __launch_bounds__(kBlock) void bwdKernel(Dev d, const int2* __restrict__ items) {
  extern __shared__ __align__(16) char lds[];
  const int2 it = items[blockIdx.x];
  const int node = it.x;
  switch (d.kind[node]) {
    case PPSG_DECIDE:
      switch (d.na[node]) {
        case 1: decideTile<TR, MODE, 1>(d, node, it.y); break;
        case 2: decideTile<TR, MODE, 2>(d, node, it.y); break;
        case 3: decideTile<TR, MODE, 3>(d, node, it.y); break;
        case 4: decideTile<TR, MODE, 4>(d, node, it.y); break;
        case 5: decideTile<TR, MODE, 5>(d, node, it.y); break;
        case 6: decideTile<TR, MODE, 6>(d, node, it.y); break;
        default: decideTile<TR, MODE, 0>(d, node, it.y); break;
      }
      break;
    case PPSG_CHANCE:
      chanceTile<TR>(d, node, it.y);
      break;
    case PPSG_FOLD:
      foldNode<TR>(d, node, lds);
      break;
    default:
      showdownNode<TR>(d, node, lds);
      break;
  }
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
  // per (tr, depth): device item list + count, and whether it has terminals
  std::vector<int2*> fwdItems[2], bwdItems[2];
  std::vector<int> fwdCount[2], bwdCount[2];
  std::vector<char> bwdTerminal[2];
  Ctl* ctl = nullptr;
  float* coef = nullptr;
  int64_t coefCap = 0;
  int64_t coefLen = 0;   // valid schedule entries
  int64_t counter = 0;   // host mirror of the device schedule index
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

int ppsg_engine::launchHalf(int tr, int mode, hipStream_t s) {
  const int fromStrat = mode == PPSG_MODE_CFR ? 0 : 1;
  for (int dpt = 0; dpt <= maxDepth; ++dpt) {
    const int cnt = fwdCount[tr][dpt];
    if (cnt == 0) continue;
    if (tr == 0)
      fwdKernel<0><<<cnt, kBlock, 0, s>>>(dev, fwdItems[0][dpt], fromStrat);
    else
      fwdKernel<1><<<cnt, kBlock, 0, s>>>(dev, fwdItems[1][dpt], fromStrat);
    HIP_TRY(hipGetLastError());
  }
  for (int dpt = maxDepth; dpt >= 0; --dpt) {
    const int cnt = bwdCount[tr][dpt];
    if (cnt == 0) continue;
    const size_t lds = bwdTerminal[tr][dpt] ? terminalLds : 0;
    const int2* it = bwdItems[tr][dpt];
#define PPSG_BWD(T, M) bwdKernel<T, M><<<cnt, kBlock, lds, s>>>(dev, it)
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
  }
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

extern "C" {

const char* ppsg_last_error(void) { return g_err.c_str(); }

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
  int maxList = 0;
  for (int p = 0; p < 2; ++p) {
    for (int c = 0; c < nc; ++c) maxList = std::max(maxList, t->cls_n[p][c]);
    if (E->uploadC(d.clsN[p], t->cls_n[p], nc) || E->uploadC(d.clsOff[p], t->cls_list_off[p], nc) ||
        E->uploadC(d.listCards[p], t->list_cards[p], t->list_total[p]) ||
        E->uploadC(d.listSame[p], t->list_same[p], t->list_total[p]) ||
        E->uploadC(d.toParent[p], t->to_parent[p], t->list_total[p]) ||
        E->uploadC(d.fromParentOff[p], t->from_parent_off[p], nc) ||
        E->uploadC(d.fromParent[p], t->from_parent[p], std::max<int64_t>(t->from_parent_total[p], 1)) ||
        E->uploadC(d.segOff[p], t->seg_off[p], (int64_t)nc * 53) ||
        E->uploadC(d.segSlot[p], t->seg_slot[p], 2 * t->list_total[p]) ||
        E->uploadC(d.sorted[p], t->sorted[p], t->list_total[p]) ||
        E->uploadC(d.sdMeta[p], t->sd_meta[p], 4 * t->list_total[p]) ||
        E->uploadC(d.reachOff[p], t->reach_off[p], t->n_nodes))
      return bail(-1);
  }
  if (maxList > 65535) return bail(fail("combo list too long"));
  const int64_t N = t->n_nodes;
  if (E->uploadC(d.clsBoardN, t->cls_board_n, nc) || E->uploadC(d.kind, t->kind, N) ||
      E->uploadC(d.owner, t->owner, N) || E->uploadC(d.na, t->na, N) ||
      E->uploadC(d.cls, t->cls, N) || E->uploadC(d.childBase, t->child_base, N) ||
      E->uploadC(d.childFlat, t->child_flat, std::max<int64_t>(t->n_child, 1)) ||
      E->uploadC(d.regOff, t->reg_off, N) || E->uploadC(d.cfvOff, t->cfv_off, N) ||
      E->uploadC(d.termVal, t->term_val, N * 8))
    return bail(-1);
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
    if (hipMemcpy(d.reach[tr], t->root_reach[tr], sizeof(float) * t->n_base[1 - tr],
                  hipMemcpyHostToDevice) != hipSuccess)
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
  // LDS for one terminal workgroup: raw reach (f32) + 3n+1 f64 prefix slots
  E->terminalLds = (size_t)((maxList * 4 + 15) & ~15) + (size_t)(3 * maxList + 1) * 8;
  if (E->terminalLds > 160 * 1024) return bail(fail("terminal LDS too large"));

  // work lists
  const int D = t->max_depth + 1;
  std::vector<std::vector<int2>> fwd[2], bwd[2];
  for (int tr = 0; tr < 2; ++tr) {
    fwd[tr].assign(D, {});
    bwd[tr].assign(D, {});
    E->bwdTerminal[tr].assign(D, 0);
  }
  auto tiles = [](int n) { return (n + kTile - 1) / kTile; };
  for (int64_t u = 0; u < N; ++u) {
    const int dep = t->depth[u], cl = t->cls[u];
    const int k = t->kind[u];
    for (int tr = 0; tr < 2; ++tr) {
      const int op = 1 - tr;
      if (k == PPSG_DECIDE) {
        if (t->owner[u] == op)
          for (int q = 0; q < tiles(t->cls_n[op][cl]); ++q) fwd[tr][dep].push_back({(int)u, q});
        for (int q = 0; q < tiles(t->cls_n[tr][cl]); ++q) bwd[tr][dep].push_back({(int)u, q});
      } else if (k == PPSG_CHANCE) {
        const int nb = t->na[u];
        if (nb > kMaxBranch) return bail(fail("too many chance branches"));
        for (int b = 0; b < nb; ++b) {
          const int ch = t->child_flat[t->child_base[u] + b];
          for (int q = 0; q < tiles(t->cls_n[op][t->cls[ch]]); ++q)
            fwd[tr][dep].push_back({(int)u, (b << 16) | q});
        }
        for (int q = 0; q < tiles(t->cls_n[tr][cl]); ++q) bwd[tr][dep].push_back({(int)u, q});
      } else {
        bwd[tr][dep].push_back({(int)u, 0});
        E->bwdTerminal[tr][dep] = 1;
      }
    }
  }
  for (int tr = 0; tr < 2; ++tr) {
    E->fwdItems[tr].assign(D, nullptr);
    E->bwdItems[tr].assign(D, nullptr);
    E->fwdCount[tr].assign(D, 0);
    E->bwdCount[tr].assign(D, 0);
    for (int dep = 0; dep < D; ++dep) {
      // terminal workgroups first: they are the long ones
      std::stable_sort(bwd[tr][dep].begin(), bwd[tr][dep].end(), [&](int2 a, int2 b) {
        const bool ta = t->kind[a.x] == PPSG_FOLD || t->kind[a.x] == PPSG_SHOWDOWN;
        const bool tb = t->kind[b.x] == PPSG_FOLD || t->kind[b.x] == PPSG_SHOWDOWN;
        return ta && !tb;
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
  // Algorithmic bytes of one iteration, split by pass: every f32 row
  // element read or written once per use (index tables excluded).
  for (int64_t u = 0; u < N; ++u) {
    const int cl = t->cls[u];
    const double n0 = t->cls_n[0][cl], n1 = t->cls_n[1][cl];
    if (t->kind[u] == PPSG_DECIDE) {
      const double a = t->na[u];
      const double np = t->owner[u] == 0 ? n0 : n1, nq = t->owner[u] == 0 ? n1 : n0;
      E->fwdBytes += 4 * np * (2 * a + 1);
      E->bwdBytes += 4 * np * (5 * a + 1) + 4 * nq * (a + 1);
    } else if (t->kind[u] == PPSG_CHANCE) {
      for (int b = 0; b < t->na[u]; ++b) {
        const int ccl = t->cls[t->child_flat[t->child_base[u] + b]];
        const double nc = (double)t->cls_n[0][ccl] + t->cls_n[1][ccl];
        E->fwdBytes += 8 * nc;
        E->bwdBytes += 4 * nc;
      }
      E->bwdBytes += 4 * (n0 + n1);
    } else {
      E->bwdBytes += 8 * (n0 + n1);
    }
  }
  if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&E->ev0) != hipSuccess || hipEventCreate(&E->ev1) != hipSuccess)
    return bail(fail("stream/event creation failed"));
  // terminal kernels may need more than the default dynamic LDS
  if (E->terminalLds > 64 * 1024) {
    const void* ks[] = {(const void*)bwdKernel<0, 0>, (const void*)bwdKernel<0, 1>,
                        (const void*)bwdKernel<0, 2>, (const void*)bwdKernel<1, 0>,
                        (const void*)bwdKernel<1, 1>, (const void*)bwdKernel<1, 2>};
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

// Every launch reads coef[counter]: refuse work that would run past the table.
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
  double acc[5] = {0, 0, 0, 0, 0};
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  auto timed = [&](auto&& launch, int slot) -> int {
    HIP_TRY(hipEventRecord(a, h->stream));
    launch();
    HIP_TRY(hipGetLastError());
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
        const int cnt = h->fwdCount[tr][dpt];
        if (!cnt) continue;
        const int2* it2 = h->fwdItems[tr][dpt];
        rc = timed([&] {
          if (tr == 0) fwdKernel<0><<<cnt, kBlock, 0, h->stream>>>(h->dev, it2, 0);
          else fwdKernel<1><<<cnt, kBlock, 0, h->stream>>>(h->dev, it2, 0);
        }, 0);
      }
      for (int dpt = h->maxDepth; dpt >= 0 && rc == 0; --dpt) {
        const int cnt = h->bwdCount[tr][dpt];
        if (!cnt) continue;
        const size_t lds = h->bwdTerminal[tr][dpt] ? h->terminalLds : 0;
        const int2* it2 = h->bwdItems[tr][dpt];
        rc = timed([&] {
          if (tr == 0) bwdKernel<0, PPSG_MODE_CFR><<<cnt, kBlock, lds, h->stream>>>(h->dev, it2);
          else bwdKernel<1, PPSG_MODE_CFR><<<cnt, kBlock, lds, h->stream>>>(h->dev, it2);
        }, 2);
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
  for (int i = 0; i < cap && i < 5; ++i) out[i] = acc[i];
  return 0;
}

}  // extern "C"

// serve.hip — continuous-batching projection server over resident octrees.
//
// The reference has no server: its harness projects one query batch at a time with blocking
// calls (scene.cpp:916-966).  Here clients submit projection requests (tree, kernel,
// parameters, queries, n_iter) from any thread, and the server schedules them a projection step
// at a time:
//
// * Queueing.  A request is cut into chunks of at most `chunk` queries.  Interactive requests
//   (at most kInteractivePts projected points) go ahead of bulk ones; within a class the
//   earliest deadline goes first (arrival + 1 ms per kNominalPtsPerMs projected points), then
//   arrival order.  Bulk work gets a launch at least every kBulkMaxWaitMs, so interactive
//   requests cannot starve it.
// * Residency.  Chunks taken by a step launch are staged into a device query pool (pinned host
//   staging -> device; Morton key in the tree's box; radix sort; gather into pool blocks of
//   kBatchWG slots) and stay there as a run until their last step; consecutive chunks of one
//   request taken together form one run, sorted and walked together.
// * Steps.  Each of `inflight` streams keeps one step launch on the device; when it completes,
//   the worker fills the stream's next launch with up to `max_batch_queries` queries of the most
//   urgent runs, resident or queued, of one class.  Bulk launches take kBulkSteps steps per run,
//   so bulk work yields to interactive work every few milliseconds; interactive launches take up
//   to kInteractiveSteps (a whole request, at the usual step counts) and raise their
//   wavefronts' priority.  A request that arrives while others are on the GPU therefore starts
//   at the next launch boundary of any stream; a run not picked for a launch waits in the pool
//   at no cost.
// * Hand-back.  A run taking its last steps writes its results straight to the stream's pinned,
//   host-coherent staging (rows in submission order) and flags each of its workgroups done; the
//   worker polls the flags and copies a run to its caller as soon as the run is done, without
//   waiting for the rest of its launch.
//
// * Admission by memory footprint.  A run holds pool blocks (kBatchWG query slots of 36 B, 68 B
//   in the FP64 form) from its first launch to its hand-back; chunks are admitted while the
//   pool has blocks for them and wait in their queue otherwise.  mlod_server_set_budget caps the
//   HBM the server holds with its trees (resident trees + staging + pool): the pool is resized to
//   what the budget leaves, trees that would exceed it are refused, and so are requests whose
//   chunks could never be resident under it.
// * Cancellation.  mlod_server_cancel drops a request's queued chunks and resident runs at the
//   next launch boundary; runs already in a launch finish it and are discarded.  The request
//   completes with status 1 and its output buffers hold whatever runs were handed back before.
//
// A query gives the bits of a standalone projection whatever the mix and the split of its steps
// over launches: a lane's traversal and fit depend only on its query and tree, and the position
// carried from one step to the next is the one the next step starts from (the float position,
// or the FP64 one in the FP64 form).  A request's tree is resolved when it is submitted.  Only the
// queues, the tree list, the ticket table and the completion list are shared with callers (under
// the mutex); the pool, runs and streams belong to the worker thread, and staging copies and
// launches run outside the mutex.
#include <cmath>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <map>
#include <mutex>
#include <queue>
#include <thread>
#include <vector>

#include "mlod_internal.h"

namespace mlod {
namespace {

using clk = std::chrono::steady_clock;
constexpr int BS = 256;
constexpr uint32_t kMaxRuns = 1024;         // runs per launch: the run index fits 10 bits above the Morton code
constexpr double kNominalPtsPerMs = 2.0e4;  // deadline slack: 1 ms per 20k projected points
constexpr double kInteractivePts = 2.0e6;   // interactive class: at most 2M projected points (200k queries x 10 steps)
constexpr uint32_t kInteractiveSteps = 16;  // steps per launch for interactive runs (all, for n_iter <= 16)
constexpr uint32_t kBulkSteps = 2;          // steps per launch for bulk runs
constexpr uint32_t kPoolLaunches = 4;       // pool capacity, in launches' worth of blocks per stream
constexpr double kBulkMaxWaitMs = 20.0;     // bulk work gets a launch at least this often

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// key = run index << 30 | Morton code of the query in its tree's box, over the staged queries of
// the launch's new runs (reqs[0..nReq), staged back to back from their slot0)
__global__ __launch_bounds__(BS) void k_batch_keys(const float* __restrict__ in, uint32_t slots,
                                                   const ReqDesc* __restrict__ reqs, uint32_t nReq,
                                                   uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= slots) return;
  uint32_t lo = 0, hi = nReq;  // last run with slot0 <= i
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (reqs[mid].slot0 <= i)
      lo = mid;
    else
      hi = mid;
  }
  const ReqDesc& R = reqs[lo];
  uint32_t c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float u = (in[3 * (uint64_t)i + a] - R.org[a]) * R.scale[a];
    u = fminf(fmaxf(u, 0.f), 1023.f);
    c[a] = (uint32_t)u;
  }
  key[i] = ((uint64_t)lo << 30) | (spread10(c[0]) | (spread10(c[1]) << 1) | (spread10(c[2]) << 2));
  idx[i] = i;
}

// sorted staged query j -> its pool slot, recording its index within the run (and, in the FP64
// form, its FP64 position)
__global__ __launch_bounds__(BS) void k_pool_gather(const float* __restrict__ in, const uint64_t* __restrict__ key,
                                                    const uint32_t* __restrict__ perm, uint32_t slots,
                                                    const ReqDesc* __restrict__ reqs, float4* __restrict__ q,
                                                    float4* __restrict__ qn, double4* __restrict__ qd,
                                                    uint32_t* __restrict__ pidx) {
  const uint32_t j = blockIdx.x * BS + threadIdx.x;
  if (j >= slots) return;
  const uint32_t s = perm[j];
  const ReqDesc& R = reqs[(uint32_t)(key[j] >> 30)];
  const uint32_t d = R.pad0 + (j - R.slot0);
  const float x = in[3 * (uint64_t)s], y = in[3 * (uint64_t)s + 1], z = in[3 * (uint64_t)s + 2];
  q[d] = make_float4(x, y, z, 0.f);
  qn[d] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (qd) qd[d] = make_double4(x, y, z, 0.0);
  pidx[d] = s - R.slot0;
}

// One request as submitted.
struct Request {
  uint64_t ticket;
  const mlod_tree* tree;  // resolved at submission
  mlod_kernel kernel;
  float cutoff;
  uint32_t minDepth;
  uint32_t nIter;
  float stepMax;
  const float* in;
  float* outP;
  float* outN;
  uint32_t m;
  uint32_t left;  // queries not yet handed back (worker thread only)
  bool interactive;
  bool started;
  std::atomic<bool> cancelled{false};
  double deadline;  // ms since server creation
  clk::time_point submitted, launched;
};

// Queries [off, off + m) of a request, queued.
struct Chunk {
  Request* req;
  uint32_t off, m;
  uint64_t seq;  // submission order
};

struct ChunkLater {
  bool operator()(const Chunk& a, const Chunk& b) const {
    return a.req->deadline != b.req->deadline ? a.req->deadline > b.req->deadline : a.seq > b.seq;
  }
};
using ChunkQueue = std::priority_queue<Chunk, std::vector<Chunk>, ChunkLater>;

// Queries [off, off + m) of a request, resident in pool blocks [blk0, blk0 + nblk).
struct Run {
  Request* req;
  uint32_t off, m;
  uint64_t seq;
  uint32_t blk0, nblk;
  uint32_t stepsLeft;
  uint32_t steps;  // steps in the launch that holds it
  int stream;      // stream whose launch holds it, -1 while it waits in the pool
  bool fresh;      // not staged yet: its launch stages it first
  bool handed;     // results handed back (its launch may still run others)
};

struct Stream {
  hipStream_t s = nullptr;
  hipEvent_t ev = nullptr;
  // staging of new runs
  float* hIn = nullptr;
  float* dIn = nullptr;
  uint64_t* dKey = nullptr;
  uint64_t* dKey2 = nullptr;
  uint32_t* dIdx = nullptr;
  uint32_t* dPerm = nullptr;
  void* sortTmp = nullptr;
  size_t sortBytes = 0;
  // launch tables
  ReqDesc* hReq = nullptr;
  ReqDesc* dReq = nullptr;
  uint2* hBlk = nullptr;
  uint2* dBlk = nullptr;
  // results of the runs taking their last steps, written by the launch (pinned, host-coherent):
  // rows [0, cap) positions, [cap, 2 cap) normals; flag[launch block] = seq when the block is done
  float* hOut = nullptr;
  float* dOut = nullptr;  // device address of hOut
  uint32_t* hFlag = nullptr;
  uint32_t* dFlag = nullptr;
  uint32_t seq = 0;
  std::vector<Run*> runs;         // runs of the launch
  std::vector<uint32_t> outOff;   // per run: its rows in hOut, kNoOut if it is not handed back
  std::vector<uint32_t> blk0;     // per run: its first launch block
  std::vector<uint32_t> seen;     // per run: launch blocks seen done so far
  bool busy = false;
};

}  // namespace
}  // namespace mlod

struct mlod_server {
  uint32_t cap = 0;  // queries per launch
  int semantics = MLOD_SEM_REF_GPU;
  int precision = MLOD_PREC_FP32C;
  int device = 0;
  std::vector<mlod_tree*> trees;
  std::map<uint64_t, mlod::Request*> live;  // outstanding requests by ticket
  uint32_t chunk = 0;
  // memory footprint (bytes of device memory): budget (0 = none), registered trees, staging of
  // all streams, pool; a budget change is applied by the worker when it is idle
  uint64_t budget = 0, treeBytes = 0, stageBytes = 0, poolBytes = 0;
  uint64_t budgetReq = 0;
  bool budgetPending = false;
  int budgetResult = MLOD_OK;
  std::string budgetMsg;
  mlod::ChunkQueue qInt, qBulk;  // queued chunks per class
  uint64_t nextSeq = 0;
  std::vector<mlod_server_done> done;
  std::mutex mu;
  std::condition_variable cv, cvDone;
  std::thread worker;
  std::atomic<bool> stop{false};
  std::atomic<int> error{0};
  std::string errmsg;
  uint64_t nextTicket = 1;
  uint64_t launches = 0, queries = 0, requests = 0;
  mlod::clk::time_point t0 = mlod::clk::now();
  // worker thread only
  double lastBulkMs = 0.0;
  std::vector<mlod::Stream> streams;
  std::vector<mlod::Run*> runs;  // resident runs
  float4* dQ = nullptr;          // query pool
  float4* dQn = nullptr;
  double4* dQd = nullptr;        // FP64 positions (FP64 form only)
  uint32_t* dPidx = nullptr;
  uint32_t poolBlocks = 0, freeBlocks = 0;
  std::map<uint32_t, uint32_t> freeRanges;  // first block -> length
};

namespace mlod {
namespace {

template <class T>
hipError_t dmalloc(T** p, size_t n) {
  return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}
template <class T>
hipError_t hmalloc(T** p, size_t n) {
  return hipHostMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault);
}

inline uint32_t blocks_of(uint32_t m) { return (m + kBatchWG - 1) / kBatchWG; }
// blocks of one launch: every run wastes less than one block
inline uint32_t max_launch_blocks(uint32_t cap) { return blocks_of(cap) + kMaxRuns; }

void free_stream(Stream& b) {
  if (b.s) (void)hipStreamSynchronize(b.s);
  void* d[] = {b.dIn, b.dKey, b.dKey2, b.dIdx, b.dPerm, b.sortTmp, b.dReq, b.dBlk};
  for (void* p : d)
    if (p) (void)hipFree(p);
  void* h[] = {b.hIn, b.hReq, b.hBlk, b.hOut, b.hFlag};
  for (void* p : h)
    if (p) (void)hipHostFree(p);
  if (b.ev) (void)hipEventDestroy(b.ev);
  if (b.s) (void)hipStreamDestroy(b.s);
  b = Stream{};
}

// device bytes of one stream's staging and launch tables
uint64_t stream_device_bytes(uint32_t cap, size_t sortBytes) {
  const uint64_t nblk = max_launch_blocks(cap);
  return 3ull * cap * sizeof(float) + 2ull * cap * sizeof(uint64_t) + 2ull * cap * sizeof(uint32_t) +
         (uint64_t)kMaxRuns * sizeof(ReqDesc) + nblk * sizeof(uint2) + sortBytes + 16;
}

// device bytes of one pool block
uint64_t block_bytes(int precision) {
  return (uint64_t)kBatchWG * (2 * sizeof(float4) + sizeof(uint32_t) + (precision == MLOD_PREC_FP64 ? sizeof(double4) : 0));
}

hipError_t alloc_stream(Stream& b, uint32_t cap) {
  hipError_t e;
  const size_t nblk = max_launch_blocks(cap);
  if ((e = hipStreamCreate(&b.s)) || (e = hipEventCreateWithFlags(&b.ev, hipEventDisableTiming))) return e;
  if ((e = hmalloc(&b.hIn, 3 * (size_t)cap)) || (e = hmalloc(&b.hReq, kMaxRuns)) || (e = hmalloc(&b.hBlk, nblk)))
    return e;
  if ((e = hipHostMalloc((void**)&b.hOut, 6 * (size_t)cap * sizeof(float), hipHostMallocCoherent)) ||
      (e = hipHostMalloc((void**)&b.hFlag, nblk * sizeof(uint32_t), hipHostMallocCoherent)))
    return e;
  std::memset(b.hFlag, 0, nblk * sizeof(uint32_t));
  if ((e = hipHostGetDevicePointer((void**)&b.dOut, b.hOut, 0)) ||
      (e = hipHostGetDevicePointer((void**)&b.dFlag, b.hFlag, 0)))
    return e;
  if ((e = dmalloc(&b.dIn, 3 * (size_t)cap)) || (e = dmalloc(&b.dKey, cap)) || (e = dmalloc(&b.dKey2, cap)) ||
      (e = dmalloc(&b.dIdx, cap)) || (e = dmalloc(&b.dPerm, cap)) || (e = dmalloc(&b.dReq, kMaxRuns)) ||
      (e = dmalloc(&b.dBlk, nblk)))
    return e;
  if ((e = rocprim::radix_sort_pairs(nullptr, b.sortBytes, b.dKey, b.dKey2, b.dIdx, b.dPerm, cap, 0, 40, b.s))) return e;
  return hipMalloc(&b.sortTmp, b.sortBytes + 16);
}

void free_pool(mlod_server* S) {
  void* d[] = {S->dQ, S->dQn, S->dQd, S->dPidx};
  for (void* p : d)
    if (p) (void)hipFree(p);
  S->dQ = S->dQn = nullptr;
  S->dQd = nullptr;
  S->dPidx = nullptr;
  S->poolBytes = 0;
  S->poolBlocks = S->freeBlocks = 0;
  S->freeRanges.clear();
}

hipError_t alloc_pool(mlod_server* S, uint32_t blocks) {
  hipError_t e;
  const size_t slots = (size_t)blocks * kBatchWG;
  if ((e = dmalloc(&S->dQ, slots)) || (e = dmalloc(&S->dQn, slots)) || (e = dmalloc(&S->dPidx, slots))) return e;
  if (S->precision == MLOD_PREC_FP64 && (e = dmalloc(&S->dQd, slots))) return e;
  S->poolBlocks = S->freeBlocks = blocks;
  S->poolBytes = blocks * block_bytes(S->precision);
  S->freeRanges.clear();
  S->freeRanges[0] = blocks;
  return hipSuccess;
}

// Blocks a bulk run must leave free for interactive ones: one launch's worth, at most half the
// pool (so a small, budgeted pool still admits bulk chunks).
uint32_t bulk_reserve(const mlod_server* S) { return std::min(max_launch_blocks(S->cap), S->poolBlocks / 2); }

// First fit; false when no free range is long enough.
bool pool_alloc(mlod_server* S, uint32_t n, uint32_t& blk0) {
  for (auto it = S->freeRanges.begin(); it != S->freeRanges.end(); ++it) {
    if (it->second < n) continue;
    blk0 = it->first;
    const uint32_t rest = it->second - n;
    S->freeRanges.erase(it);
    if (rest) S->freeRanges[blk0 + n] = rest;
    S->freeBlocks -= n;
    return true;
  }
  return false;
}

void pool_free(mlod_server* S, uint32_t blk0, uint32_t n) {
  auto it = S->freeRanges.emplace(blk0, n).first;
  auto nx = std::next(it);
  if (nx != S->freeRanges.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    S->freeRanges.erase(nx);
  }
  if (it != S->freeRanges.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      S->freeRanges.erase(it);
    }
  }
  S->freeBlocks += n;
}

double ms_since(const mlod_server* S, clk::time_point t) {
  return std::chrono::duration<double, std::milli>(t - S->t0).count();
}

// Scheduling order: class (interactive first), deadline, submission.
struct Urgency {
  int cls;
  double deadline;
  uint64_t seq;
  bool operator<(const Urgency& o) const {
    if (cls != o.cls) return cls < o.cls;
    if (deadline != o.deadline) return deadline < o.deadline;
    return seq < o.seq;
  }
};
inline Urgency urgency(const Request* q, uint64_t seq) { return Urgency{q->interactive ? 0 : 1, q->deadline, seq}; }

// A candidate of the next launch: a resident run waiting in the pool, or a queued chunk.
struct Cand {
  Urgency u;
  Run* run;
  Chunk ch;
};

// Fills st.runs with the runs of the stream's next launch: candidates of class cls, most urgent
// first, at most S->cap queries; queued chunks become new runs (consecutive chunks of a request
// one run) while the pool has free blocks for them (bulk chunks leave bulk_reserve blocks to
// interactive ones); the others stay queued.  Returns the queries picked; the sequence numbers of
// the chunks placed go to `placed`.
uint32_t take_class(mlod_server* S, Stream& st, const std::vector<Cand>& cands, int cls, std::vector<uint64_t>& placed) {
  const uint32_t steps = cls == 0 ? kInteractiveSteps : kBulkSteps;
  const uint32_t reserve = cls == 0 ? 0u : bulk_reserve(S);
  uint32_t avail = S->freeBlocks > reserve ? S->freeBlocks - reserve : 0u;
  uint32_t used = 0;
  std::vector<Chunk> taken;
  for (const Cand& c : cands) {
    const uint32_t m = c.run ? c.run->m : c.ch.m;
    if (c.u.cls != cls || used + m > S->cap || st.runs.size() >= kMaxRuns) continue;
    if (c.run) {
      used += m;
      c.run->steps = std::min(steps, c.run->stepsLeft);
      st.runs.push_back(c.run);
      continue;
    }
    Run* last = st.runs.empty() ? nullptr : st.runs.back();
    const bool merge = last && last->fresh && last->req == c.ch.req && last->off + last->m == c.ch.off;
    const uint32_t nb = merge ? blocks_of(last->m + m) - blocks_of(last->m) : blocks_of(m);
    if (nb > avail) continue;  // not admitted now: stays queued
    avail -= nb;
    used += m;
    taken.push_back(c.ch);
    if (merge) {
      last->m += m;
      continue;
    }
    Run* r = new Run{c.ch.req, c.ch.off, m, c.ch.seq, 0, 0, c.ch.req->nIter, 0, -1, true, false};
    r->steps = std::min(steps, r->stepsLeft);
    st.runs.push_back(r);
  }
  // pool blocks for the new runs; a run without a free range long enough stays queued as its chunks
  std::vector<Run*> keep;
  uint32_t picked = 0;
  for (Run* r : st.runs) {
    if (r->fresh) {
      const uint32_t nb = blocks_of(r->m);
      uint32_t b0 = 0;
      if (!pool_alloc(S, nb, b0)) {
        delete r;
        continue;
      }
      r->blk0 = b0;
      r->nblk = nb;
      S->runs.push_back(r);
      for (const Chunk& c : taken)
        if (c.req == r->req && c.off >= r->off && c.off < r->off + r->m) placed.push_back(c.seq);
    }
    keep.push_back(r);
    picked += r->m;
  }
  st.runs.swap(keep);
  return picked;
}

// Drops the cancelled work the worker holds outside launches: queued chunks (from `cands`) and
// resident runs waiting in the pool.  Their queries count as handed back; a cancelled request
// whose queries are all accounted for completes (status 1).
void drop_cancelled(mlod_server* S, std::vector<Cand>& cands, std::vector<mlod_server_done>& out,
                    std::vector<Request*>& finished) {
  const auto now = clk::now();
  auto account = [&](Request* q, uint32_t m) {
    q->left -= m;
    if (q->left == 0) {
      mlod_server_done d{};
      d.ticket = q->ticket;
      d.latency_ms = std::chrono::duration<double, std::milli>(now - q->submitted).count();
      d.queue_ms = std::chrono::duration<double, std::milli>(q->launched - q->submitted).count();
      d.completed_s = std::chrono::duration<double>(now - S->t0).count();
      d.status = 1;
      out.push_back(d);
      finished.push_back(q);
    }
  };
  std::vector<Cand> keep;
  for (Cand& c : cands) {
    Request* q = c.run ? c.run->req : c.ch.req;
    if (!q->cancelled.load(std::memory_order_relaxed)) {
      keep.push_back(c);
      continue;
    }
    if (c.run) {
      pool_free(S, c.run->blk0, c.run->nblk);
      S->runs.erase(std::find(S->runs.begin(), S->runs.end(), c.run));
      account(q, c.run->m);
      delete c.run;
    } else {
      account(q, c.ch.m);
    }
  }
  cands.swap(keep);
}

// Picks the next launch of stream st; returns its queries (0: nothing to launch).  Interactive
// work goes first, except that bulk work waiting for more than kBulkMaxWaitMs since the last
// bulk launch gets the launch, so a stream of interactive requests cannot starve it.
uint32_t pick_runs(mlod_server* S, Stream& st, bool& interactive, std::vector<mlod_server_done>& dropped,
                   std::vector<Request*>& finished) {
  const double now = ms_since(S, clk::now());
  std::vector<Cand> cands;
  for (Run* r : S->runs)
    if (r->stream < 0) cands.push_back(Cand{urgency(r->req, r->seq), r, Chunk{}});
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (ChunkQueue* Q : {&S->qInt, &S->qBulk}) {
      uint32_t got = 0;
      while (!Q->empty() && got < S->cap) {
        const Chunk c = Q->top();
        Q->pop();
        got += c.m;
        cands.push_back(Cand{urgency(c.req, c.seq), nullptr, c});
      }
    }
  }
  drop_cancelled(S, cands, dropped, finished);
  if (cands.empty()) return 0;
  std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.u < b.u; });
  int cls = cands[0].u.cls;
  if (cls == 0 && cands.back().u.cls == 1 && now - S->lastBulkMs > kBulkMaxWaitMs) cls = 1;
  std::vector<uint64_t> placed;
  uint32_t picked = take_class(S, st, cands, cls, placed);
  if (picked == 0 && cands.front().u.cls != cands.back().u.cls) {
    // nothing of that class fits the pool now: the other class
    cls = 1 - cls;
    picked = take_class(S, st, cands, cls, placed);
  }
  interactive = cls == 0;
  if (picked && !interactive) S->lastBulkMs = now;
  std::sort(placed.begin(), placed.end());
  std::lock_guard<std::mutex> lk(S->mu);
  for (const Cand& c : cands)
    if (!c.run && !std::binary_search(placed.begin(), placed.end(), c.ch.seq))
      (c.ch.req->interactive ? S->qInt : S->qBulk).push(c.ch);
  return picked;
}

// Stages the new runs of st.runs and launches the steps of all of them (worker thread, without
// the mutex).  Runs taking their last steps get result rows in the stream's host staging.
int launch_step(mlod_server* S, Stream& st, bool interactive) {
  const int sid = (int)(&st - S->streams.data());
  // new runs first: their staging keys index them
  std::stable_partition(st.runs.begin(), st.runs.end(), [](const Run* r) { return r->fresh; });
  uint32_t nNew = 0, staged = 0, nBlk = 0, outRows = 0;
  const size_t nRuns = st.runs.size();
  st.outOff.assign(nRuns, kNoOut);
  st.blk0.assign(nRuns, 0);
  st.seen.assign(nRuns, 0);
  const auto now = clk::now();
  for (size_t k = 0; k < nRuns; ++k) {
    Run* r = st.runs[k];
    r->stream = sid;
    Request& q = *r->req;
    const mlod_tree* t = q.tree;
    ReqDesc& d = st.hReq[k];
    d.tv = t->view();
    d.kp = make_kparams(q.kernel);
    d.cutoff = q.cutoff;
    d.stepMax = q.stepMax < 0.f ? t->radius : q.stepMax;
    d.minDepth = q.minDepth;
    d.nIter = r->steps;
    d.count = r->m;
    d.pad0 = r->blk0 * kBatchWG;
    d.slot0 = 0;
    d.outOff = kNoOut;
    for (int a = 0; a < 3; ++a) {
      const float ext = t->BB[a] - t->bb[a];
      d.org[a] = t->bb[a];
      d.scale[a] = ext > 0.f ? 1024.f / ext : 0.f;
    }
    if (r->fresh) {
      d.slot0 = staged;
      std::memcpy(st.hIn + 3 * (size_t)staged, q.in + 3 * (size_t)r->off, (size_t)r->m * 3 * sizeof(float));
      staged += r->m;
      ++nNew;
      if (!q.started) {
        q.started = true;
        q.launched = now;
      }
    }
    st.blk0[k] = nBlk;
    for (uint32_t b = 0; b < r->nblk; ++b) st.hBlk[nBlk++] = make_uint2((uint32_t)k, r->blk0 + b);
    if (r->steps == r->stepsLeft) {
      d.outOff = st.outOff[k] = outRows;
      outRows += r->m;
    }
  }
  hipError_t e = hipMemcpyAsync(st.dReq, st.hReq, nRuns * sizeof(ReqDesc), hipMemcpyHostToDevice, st.s);
  if (!e) e = hipMemcpyAsync(st.dBlk, st.hBlk, nBlk * sizeof(uint2), hipMemcpyHostToDevice, st.s);
  if (!e && staged) {
    const uint32_t grid = (staged + BS - 1) / BS;
    int runBits = 1;
    while ((1u << runBits) < nNew) ++runBits;
    e = hipMemcpyAsync(st.dIn, st.hIn, (size_t)staged * 3 * sizeof(float), hipMemcpyHostToDevice, st.s);
    if (!e) {
      k_batch_keys<<<grid, BS, 0, st.s>>>(st.dIn, staged, st.dReq, nNew, st.dKey, st.dIdx);
      e = hipGetLastError();
    }
    if (!e)
      e = rocprim::radix_sort_pairs(st.sortTmp, st.sortBytes, st.dKey, st.dKey2, st.dIdx, st.dPerm, staged, 0,
                                    30 + runBits, st.s);
    if (!e) {
      k_pool_gather<<<grid, BS, 0, st.s>>>(st.dIn, st.dKey2, st.dPerm, staged, st.dReq, S->dQ, S->dQn, S->dQd,
                                           S->dPidx);
      e = hipGetLastError();
    }
  }
  if (e) return hip_fail(e, "serving launch", __FILE__, __LINE__);
  for (Run* r : st.runs) r->fresh = false;
  st.seq += 1;
  const BatchOut bo{S->dPidx, st.dOut, S->cap, st.dFlag, st.seq};
  if (int rc = launch_project_batch(st.dReq, st.dBlk, nBlk, S->dQ, S->dQn, S->dQd, bo, interactive ? 1u : 0u,
                                    S->semantics, S->precision, st.s))
    return rc;
  e = hipEventRecord(st.ev, st.s);
  if (e) return hip_fail(e, "serving launch", __FILE__, __LINE__);
  return MLOD_OK;
}

// Copies the results of run k of st's launch (all its workgroups flagged done) to its caller,
// frees its pool blocks, and collects its request if that is complete.  Returns its queries.
uint64_t hand_back(mlod_server* S, Stream& st, size_t k, clk::time_point now, std::vector<mlod_server_done>& out,
                   std::vector<Request*>& finished) {
  Run* r = st.runs[k];
  Request& q = *r->req;
  const size_t o = st.outOff[k];
  const bool cancelled = q.cancelled.load(std::memory_order_relaxed);
  if (q.outP && !cancelled)
    std::memcpy(q.outP + 3 * (size_t)r->off, st.hOut + 3 * o, (size_t)r->m * 3 * sizeof(float));
  if (q.outN && !cancelled)
    std::memcpy(q.outN + 3 * (size_t)r->off, st.hOut + 3 * ((size_t)S->cap + o), (size_t)r->m * 3 * sizeof(float));
  pool_free(S, r->blk0, r->nblk);
  S->runs.erase(std::find(S->runs.begin(), S->runs.end(), r));
  r->handed = true;
  q.left -= r->m;
  if (q.left == 0) {
    mlod_server_done d{};
    d.ticket = q.ticket;
    d.status = cancelled ? 1 : 0;
    d.latency_ms = std::chrono::duration<double, std::milli>(now - q.submitted).count();
    d.queue_ms = std::chrono::duration<double, std::milli>(q.launched - q.submitted).count();
    d.completed_s = std::chrono::duration<double>(now - S->t0).count();
    out.push_back(d);
    finished.push_back(&q);
  }
  return r->m;
}

// Hands back the runs of st's launch whose workgroups have all flagged done.
uint64_t poll_runs(mlod_server* S, Stream& st, std::vector<mlod_server_done>& out, std::vector<Request*>& finished) {
  uint64_t handed = 0;
  const volatile uint32_t* flag = st.hFlag;
  for (size_t k = 0; k < st.runs.size(); ++k) {
    Run* r = st.runs[k];
    if (st.outOff[k] == kNoOut || r->handed) continue;
    while (st.seen[k] < r->nblk && flag[st.blk0[k] + st.seen[k]] == st.seq) ++st.seen[k];
    if (st.seen[k] < r->nblk) continue;
    std::atomic_thread_fence(std::memory_order_acquire);
    handed += hand_back(S, st, k, clk::now(), out, finished);
  }
  return handed;
}

// A stream's launch has completed: hands back the runs that took their last steps and are not
// handed back yet; the other runs wait in the pool again (worker thread, without the mutex).
uint64_t complete_step(mlod_server* S, Stream& st, std::vector<mlod_server_done>& out,
                       std::vector<Request*>& finished) {
  const auto now = clk::now();
  uint64_t handed = 0;
  for (size_t k = 0; k < st.runs.size(); ++k) {
    Run* r = st.runs[k];
    r->stepsLeft -= r->steps;
    r->stream = -1;
    if (st.outOff[k] == kNoOut && r->req->cancelled.load(std::memory_order_relaxed)) {
      st.outOff[k] = 0;  // discarded through hand_back, which copies nothing for a cancelled request
      handed += hand_back(S, st, k, now, out, finished);
      continue;
    }
    if (st.outOff[k] == kNoOut) continue;
    if (!r->handed) handed += hand_back(S, st, k, now, out, finished);
  }
  for (Run* r : st.runs)
    if (r->handed) delete r;
  st.runs.clear();
  return handed;
}

void set_error(mlod_server* S, int code, const std::string& msg) {
  std::lock_guard<std::mutex> lk(S->mu);
  if (!S->error) {
    S->errmsg = msg;
    S->error = code;
  }
  S->cvDone.notify_all();
}

// Publishes completed requests to callers (worker thread): done records, counters, and the
// requests leave the ticket table.
void publish(mlod_server* S, std::vector<mlod_server_done>& fin, std::vector<Request*>& finished, uint64_t handed) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->done.insert(S->done.end(), fin.begin(), fin.end());
    S->requests += fin.size();
    S->queries += handed;
    for (Request* r : finished) S->live.erase(r->ticket);
  }
  for (Request* r : finished) delete r;
  if (!fin.empty()) S->cvDone.notify_all();
  fin.clear();
  finished.clear();
}

// Applies a pending budget (worker thread, with nothing resident and no launch in flight): the
// pool is reallocated to the blocks the budget leaves after the trees and the staging, at most
// the default pool, at least two chunks' worth.
void apply_budget(mlod_server* S) {
  const uint64_t budget = S->budgetReq;
  const uint32_t defBlocks = kPoolLaunches * (uint32_t)S->streams.size() * max_launch_blocks(S->cap);
  uint32_t blocks = defBlocks;
  int rc = MLOD_OK;
  std::string msg;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (budget) {
      const uint64_t fixed = S->treeBytes + S->stageBytes;
      const uint64_t left = budget > fixed ? budget - fixed : 0;
      blocks = (uint32_t)std::min<uint64_t>(defBlocks, left / block_bytes(S->precision));
      if (blocks < 2 * blocks_of(S->chunk)) {
        rc = MLOD_ERR_OOM;
        msg = "admission: the budget of " + std::to_string(budget) + " bytes leaves " + std::to_string(left) +
              " bytes after the trees (" + std::to_string(S->treeBytes) + ") and staging (" +
              std::to_string(S->stageBytes) + "): less than two chunks of pool";
      }
    }
  }
  if (rc == MLOD_OK && blocks != S->poolBlocks) {
    free_pool(S);
    if (hipError_t e = alloc_pool(S, blocks)) {
      rc = MLOD_ERR_OOM;
      msg = std::string("admission: pool allocation failed: ") + hipGetErrorString(e);
      free_pool(S);
      (void)alloc_pool(S, std::min(blocks, 2 * blocks_of(S->chunk)));
    }
  }
  std::lock_guard<std::mutex> lk(S->mu);
  if (rc == MLOD_OK) S->budget = budget;
  S->budgetResult = rc;
  S->budgetMsg = msg;
  S->budgetPending = false;
  S->cvDone.notify_all();
}

void worker_main(mlod_server* S) {
  (void)hipSetDevice(S->device);
  std::vector<mlod_server_done> fin;
  std::vector<Request*> finished;
  while (true) {
    bool progressed = false, anyBusy = false;
    // completions: copy out, then publish
    for (Stream& st : S->streams) {
      if (!st.busy) continue;
      const hipError_t q = hipEventQuery(st.ev);
      uint64_t handed = 0;
      if (q == hipErrorNotReady) handed = poll_runs(S, st, fin, finished);
      if (q == hipSuccess || handed) {
        if (q == hipSuccess) {
          handed += complete_step(S, st, fin, finished);
          st.busy = false;
        } else {
          anyBusy = true;
        }
        publish(S, fin, finished, handed);
        progressed = true;
      } else if (q != hipErrorNotReady) {
        set_error(S, MLOD_ERR_HIP, std::string("serving launch failed: ") + hipGetErrorString(q));
        return;
      } else {
        anyBusy = true;
      }
    }
    // a budget change waits until nothing is resident and no launch is in flight
    bool budget = false;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      budget = S->budgetPending;
    }
    if (budget && !anyBusy && S->runs.empty()) {
      apply_budget(S);
      progressed = true;
    }
    // next launches on idle streams
    for (Stream& st : S->streams) {
      if (st.busy || budget) continue;
      bool interactive = false;
      const uint32_t picked = pick_runs(S, st, interactive, fin, finished);
      if (!fin.empty()) {
        publish(S, fin, finished, 0);
        progressed = true;
      }
      if (picked == 0) break;
      const int r = launch_step(S, st, interactive);
      if (r) {
        set_error(S, r, g_last_error());
        return;
      }
      st.busy = true;
      anyBusy = true;
      progressed = true;
      std::lock_guard<std::mutex> lk(S->mu);
      S->launches += 1;
    }
    std::unique_lock<std::mutex> lk(S->mu);
    const bool queued = !S->qInt.empty() || !S->qBulk.empty();
    if (S->stop && !queued && !anyBusy && S->runs.empty()) return;
    if (progressed) continue;
    if (anyBusy) {
      lk.unlock();
      std::this_thread::yield();
    } else {
      S->cv.wait_for(lk, std::chrono::milliseconds(1),
                     [S] { return S->stop || S->budgetPending || !S->qInt.empty() || !S->qBulk.empty(); });
    }
  }
}

}  // namespace
}  // namespace mlod

using namespace mlod;

extern "C" {

int mlod_server_create(uint32_t max_batch_queries, uint32_t chunk_queries, uint32_t inflight, int32_t semantics,
                       int32_t precision, mlod_server** out) {
  if (!out) return fail(MLOD_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (max_batch_queries < kBatchWG || max_batch_queries > (1u << 24) || inflight == 0 || inflight > 8)
    return fail(MLOD_ERR_INVALID, "bad launch capacity or inflight");
  if (chunk_queries > max_batch_queries) return fail(MLOD_ERR_INVALID, "chunk larger than a launch");
  if (precision < MLOD_PREC_FP32 || precision > MLOD_PREC_FP64) return fail(MLOD_ERR_INVALID, "bad precision");
  if (semantics != MLOD_SEM_REF_GPU && semantics != MLOD_SEM_REF_CPU) return fail(MLOD_ERR_INVALID, "bad semantics");
  mlod_server* S = new mlod_server();
  S->cap = max_batch_queries;
  S->chunk = chunk_queries ? chunk_queries : std::max<uint32_t>(kBatchWG, max_batch_queries / 4 / kBatchWG * kBatchWG);
  S->semantics = semantics;
  S->precision = precision;
  (void)hipGetDevice(&S->device);
  S->streams.resize(inflight);
  hipError_t e = alloc_pool(S, kPoolLaunches * inflight * max_launch_blocks(S->cap));
  for (size_t i = 0; !e && i < S->streams.size(); ++i) {
    e = alloc_stream(S->streams[i], S->cap);
    if (!e) S->stageBytes += stream_device_bytes(S->cap, S->streams[i].sortBytes);
  }
  if (e) {
    for (Stream& c : S->streams) free_stream(c);
    free_pool(S);
    delete S;
    return hip_fail(e, "server buffers", __FILE__, __LINE__);
  }
  S->worker = std::thread(worker_main, S);
  *out = S;
  return MLOD_OK;
}

int mlod_server_destroy(mlod_server* S) {
  if (!S) return MLOD_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  if (S->worker.joinable()) S->worker.join();
  for (Stream& st : S->streams) free_stream(st);
  // requests left behind by a failed worker
  std::vector<Request*> left;
  for (Run* r : S->runs) {
    left.push_back(r->req);
    delete r;
  }
  for (; !S->qInt.empty(); S->qInt.pop()) left.push_back(S->qInt.top().req);
  for (; !S->qBulk.empty(); S->qBulk.pop()) left.push_back(S->qBulk.top().req);
  std::sort(left.begin(), left.end());
  left.erase(std::unique(left.begin(), left.end()), left.end());
  for (Request* q : left) delete q;
  free_pool(S);
  delete S;
  return MLOD_OK;
}

int mlod_server_set_budget(mlod_server* S, uint64_t hbm_bytes) {
  if (!S) return fail(MLOD_ERR_INVALID, "NULL argument");
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->error) return fail(S->error, S->errmsg);
  if (!S->live.empty()) return fail(MLOD_ERR_STATE, "the budget changes only while no request is outstanding");
  if (hbm_bytes && S->treeBytes + S->stageBytes > hbm_bytes)
    return fail(MLOD_ERR_OOM, "admission: the registered trees (" + std::to_string(S->treeBytes) +
                                  " bytes) and staging already exceed the budget");
  S->budgetReq = hbm_bytes;
  S->budgetPending = true;
  S->cv.notify_all();
  S->cvDone.wait(lk, [S] { return !S->budgetPending || S->error; });
  if (S->error) return fail(S->error, S->errmsg);
  if (S->budgetResult) return fail(S->budgetResult, S->budgetMsg);
  return MLOD_OK;
}

int mlod_server_memory(mlod_server* S, uint64_t* out5) {
  if (!S || !out5) return fail(MLOD_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  out5[0] = S->budget;
  out5[1] = S->treeBytes;
  out5[2] = S->stageBytes;
  out5[3] = S->poolBytes;
  out5[4] = (uint64_t)S->poolBlocks * kBatchWG;
  return MLOD_OK;
}

int mlod_server_add_tree(mlod_server* S, mlod_tree* t, uint32_t* tree_id) {
  if (!S || !t || !tree_id) return fail(MLOD_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->budget && S->treeBytes + t->bytes + S->stageBytes + S->poolBytes > S->budget)
    return fail(MLOD_ERR_OOM, "admission: a tree of " + std::to_string(t->bytes) + " bytes exceeds the budget of " +
                                  std::to_string(S->budget) + " bytes (trees " + std::to_string(S->treeBytes) +
                                  ", staging and pool " + std::to_string(S->stageBytes + S->poolBytes) + ")");
  *tree_id = (uint32_t)S->trees.size();
  S->trees.push_back(t);
  S->treeBytes += t->bytes;
  return MLOD_OK;
}

int mlod_server_submit(mlod_server* S, uint32_t tree_id, const mlod_kernel* k, float cutoff, uint32_t min_depth,
                       uint32_t n_iter, float step_max, const float* queries, uint32_t m, float* out_pts, float* out_nmls,
                       uint64_t* ticket) {
  if (!S || !k || !queries || !ticket) return fail(MLOD_ERR_INVALID, "NULL argument");
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->error) return fail(S->error, S->errmsg);
  }
  if (m == 0) return fail(MLOD_ERR_INVALID, "empty request");
  if (int rc = validate_kernel(k)) return rc;
  if (!(cutoff >= 0.f)) return fail(MLOD_ERR_INVALID, "cutoff must be >= 0");
  if (std::isnan(step_max)) return fail(MLOD_ERR_INVALID, "step_max is NaN");
  Request* r = new Request();
  r->kernel = *k;
  r->cutoff = cutoff;
  r->minDepth = min_depth;
  r->nIter = n_iter;
  r->stepMax = step_max;
  r->in = queries;
  r->outP = out_pts;
  r->outN = out_nmls;
  r->m = m;
  r->left = m;
  r->started = false;
  const double pts = (double)m * std::max<uint32_t>(n_iter, 1u);
  r->interactive = pts <= kInteractivePts;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (tree_id >= S->trees.size()) {
      delete r;
      return fail(MLOD_ERR_INVALID, "unknown tree id");
    }
    // the pool must hold at least one chunk of it as a bulk run
    if (blocks_of(std::min(m, S->chunk)) + bulk_reserve(S) > S->poolBlocks) {
      delete r;
      return fail(MLOD_ERR_OOM, "admission: a chunk of this request can never be resident under the budget");
    }
    r->tree = S->trees[tree_id];
    r->ticket = S->nextTicket++;
    r->submitted = clk::now();
    r->launched = r->submitted;
    *ticket = r->ticket;
    if (n_iter == 0) {
      // no step: the queries come back unchanged with zero normals, as from a standalone projection
      if (out_pts) std::memcpy(out_pts, queries, (size_t)m * 3 * sizeof(float));
      if (out_nmls) std::memset(out_nmls, 0, (size_t)m * 3 * sizeof(float));
      mlod_server_done d{};
      d.ticket = r->ticket;
      d.latency_ms = d.queue_ms = 0.0;
      d.completed_s = std::chrono::duration<double>(clk::now() - S->t0).count();
      S->done.push_back(d);
      S->requests += 1;
      S->queries += m;
      delete r;
      S->cvDone.notify_all();
      return MLOD_OK;
    }
    r->deadline = ms_since(S, r->submitted) + pts / kNominalPtsPerMs;
    S->live[r->ticket] = r;
    ChunkQueue& Q = r->interactive ? S->qInt : S->qBulk;
    for (uint32_t off = 0; off < m; off += S->chunk) Q.push(Chunk{r, off, std::min(S->chunk, m - off), S->nextSeq++});
  }
  S->cv.notify_one();
  return MLOD_OK;
}

int mlod_server_cancel(mlod_server* S, uint64_t ticket) {
  if (!S) return fail(MLOD_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->live.find(ticket);
  if (it == S->live.end()) return fail(MLOD_ERR_INVALID, "no outstanding request with this ticket");
  it->second->cancelled.store(true, std::memory_order_relaxed);
  S->cv.notify_one();
  return MLOD_OK;
}

int mlod_server_poll(mlod_server* S, mlod_server_done* out, uint32_t cap, uint32_t* n_out, double timeout_s) {
  if (!S || !n_out) return fail(MLOD_ERR_INVALID, "NULL argument");
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->done.empty() && timeout_s > 0)
    S->cvDone.wait_for(lk, std::chrono::duration<double>(timeout_s), [S] { return !S->done.empty() || S->error; });
  if (S->error && S->done.empty()) return fail(S->error, S->errmsg);
  const uint32_t n = (uint32_t)std::min<size_t>(cap, S->done.size());
  for (uint32_t i = 0; i < n; ++i) out[i] = S->done[i];
  S->done.erase(S->done.begin(), S->done.begin() + n);
  *n_out = n;
  return MLOD_OK;
}

int mlod_server_stats(mlod_server* S, uint64_t* out4) {
  if (!S || !out4) return fail(MLOD_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  out4[0] = S->launches;
  out4[1] = S->requests;
  out4[2] = S->queries;
  out4[3] = S->qInt.size() + S->qBulk.size();
  return MLOD_OK;
}

}  // extern "C"

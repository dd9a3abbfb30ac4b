// serve.hip — continuous-batching projection server over resident octrees.
//
// The reference has no server: its harness projects one query batch at a time with blocking
// calls (scene.cpp:916-966).  Here clients submit projection requests (tree, kernel,
// parameters, queries, n_iter) from any thread.  A request is cut into chunks of at most
// `chunk` queries; a worker thread packs pending chunks into batches bounded by a slot budget
// (device memory per batch is fixed at creation) and keeps up to `inflight` batches on separate
// streams, so requests that arrive while others are on the GPU join the next batch as soon as a
// batch slot frees, and the batches' copies and kernels overlap on the device.  Chunks are taken
// earliest deadline first, the deadline being the arrival time plus 1 ms per kNominalPtsPerMs
// projected points of the request (about ten times its service time on an idle GPU): requests
// of equal size are served in arrival order, a small request that arrives behind large ones
// overtakes them, and a large one is not starved (its deadline still comes).
//
// One batch on its stream: pinned host staging -> device copy; per-slot Morton key in the
// request's tree box with the request index above it; radix sort (so each wavefront walks one
// request's neighbourhood); gather into a layout where every request starts on a workgroup
// boundary; one projection launch over all workgroups (k_project_batch: the workgroup's request
// descriptor, kernel type included, is uniform and read into scalar registers; per-query code
// and bits are those of a standalone projection); scatter back to
// submission order; device -> pinned copy; completion event.  The worker copies each request's
// results to the caller's buffers when the event has fired and stamps the completion time.
// Only the request queue and the completion list are shared with callers (under the mutex);
// staging copies and launches run outside it.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <queue>
#include <mutex>
#include <thread>
#include <vector>

#include "mlod_internal.h"

namespace mlod {
namespace {

using clk = std::chrono::steady_clock;
constexpr int BS = 256;
constexpr uint32_t kMaxReqPerBatch = 1024;  // chunk index fits 10 bits above the 30-bit Morton code
constexpr double kNominalPtsPerMs = 2.0e4;  // deadline slack: 1 ms per 20k projected points

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// key = request index << 30 | Morton code of the query in its tree's box
__global__ __launch_bounds__(BS) void k_batch_keys(const float* __restrict__ in, uint32_t slots,
                                                   const ReqDesc* __restrict__ reqs, uint32_t nReq,
                                                   uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= slots) return;
  uint32_t lo = 0, hi = nReq;  // last request with slot0 <= i
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

// sorted slot j (request key[j] >> 30) -> its place in the workgroup-aligned layout
__device__ __forceinline__ uint32_t aligned_slot(const ReqDesc* __restrict__ reqs, uint64_t key, uint32_t j) {
  const ReqDesc& R = reqs[(uint32_t)(key >> 30)];
  return R.pad0 + (j - R.slot0);
}

__global__ __launch_bounds__(BS) void k_batch_gather(const float* __restrict__ in, const uint64_t* __restrict__ key,
                                                     const uint32_t* __restrict__ perm, uint32_t slots,
                                                     const ReqDesc* __restrict__ reqs, float4* __restrict__ q,
                                                     float4* __restrict__ qn) {
  const uint32_t j = blockIdx.x * BS + threadIdx.x;
  if (j >= slots) return;
  const uint32_t s = perm[j], d = aligned_slot(reqs, key[j], j);
  q[d] = make_float4(in[3 * (uint64_t)s], in[3 * (uint64_t)s + 1], in[3 * (uint64_t)s + 2], 0.f);
  qn[d] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// out = [cap x 3 positions | cap x 3 normals] in submission order
__global__ __launch_bounds__(BS) void k_batch_scatter(const float4* __restrict__ q, const float4* __restrict__ qn,
                                                      const uint64_t* __restrict__ key, const uint32_t* __restrict__ perm,
                                                      const ReqDesc* __restrict__ reqs, uint32_t slots, uint32_t cap,
                                                      float* __restrict__ out) {
  const uint32_t j = blockIdx.x * BS + threadIdx.x;
  if (j >= slots) return;
  const uint32_t d = perm[j], a = aligned_slot(reqs, key[j], j);
  const float4 p = q[a], n = qn[a];
  out[3 * (uint64_t)d] = p.x;
  out[3 * (uint64_t)d + 1] = p.y;
  out[3 * (uint64_t)d + 2] = p.z;
  out[3 * ((uint64_t)cap + d)] = n.x;
  out[3 * ((uint64_t)cap + d) + 1] = n.y;
  out[3 * ((uint64_t)cap + d) + 2] = n.z;
}

// One request as submitted.
struct Request {
  uint64_t ticket;
  uint32_t tree;
  mlod_kernel kernel;
  float cutoff;
  uint32_t minDepth;
  uint32_t nIter;
  float stepMax;
  const float* in;
  float* outP;
  float* outN;
  uint32_t m;
  uint32_t chunksLeft;  // chunks not yet completed (worker thread only)
  clk::time_point submitted, launched;
};

// A chunk: queries [off, off + m) of a request.
struct SReq {
  Request* req;
  uint32_t off, m;
  double deadline;  // ms since server creation
  uint64_t seq;     // submission order, ties
};

struct ChunkLater {
  bool operator()(const SReq& a, const SReq& b) const {
    return a.deadline != b.deadline ? a.deadline > b.deadline : a.seq > b.seq;
  }
};

struct BatchSlot {
  hipStream_t s = nullptr;
  hipEvent_t ev = nullptr;
  float* hIn = nullptr;
  float* hOut = nullptr;
  ReqDesc* hReq = nullptr;
  float* dIn = nullptr;
  float* dOut = nullptr;
  float4* dQ = nullptr;
  float4* dQn = nullptr;
  uint64_t* dKey = nullptr;
  uint64_t* dKey2 = nullptr;
  uint32_t* dIdx = nullptr;
  uint32_t* dPerm = nullptr;
  uint32_t* dBlockReq = nullptr;   // request of each workgroup of the aligned layout
  uint32_t* hBlockReq = nullptr;
  ReqDesc* dReq = nullptr;
  void* sortTmp = nullptr;
  size_t sortBytes = 0;
  std::vector<SReq> reqs;
  std::vector<uint32_t> chunkOff;  // compact slot offset of each chunk
  uint32_t slots = 0;
  bool busy = false;
  clk::time_point launched;
};

}  // namespace
}  // namespace mlod

struct mlod_server {
  uint32_t cap = 0;
  int semantics = MLOD_SEM_REF_GPU;
  int precision = MLOD_PREC_FP32C;
  int device = 0;
  std::vector<mlod_tree*> trees;
  std::vector<mlod::BatchSlot> slots;
  uint32_t chunk = 0;
  std::priority_queue<mlod::SReq, std::vector<mlod::SReq>, mlod::ChunkLater> pending;
  uint64_t nextSeq = 0;
  std::vector<mlod_server_done> done;
  std::mutex mu;
  std::condition_variable cv, cvDone;
  std::thread worker;
  std::atomic<bool> stop{false};
  std::atomic<int> error{0};
  std::string errmsg;
  uint64_t nextTicket = 1;
  uint64_t batches = 0, queries = 0, requests = 0;
  mlod::clk::time_point t0 = mlod::clk::now();
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

void free_slot(BatchSlot& b) {
  if (b.s) (void)hipStreamSynchronize(b.s);
  void* d[] = {b.dIn, b.dOut, b.dQ, b.dQn, b.dKey, b.dKey2, b.dIdx, b.dPerm, b.dBlockReq, b.dReq, b.sortTmp};
  for (void* p : d)
    if (p) (void)hipFree(p);
  void* h[] = {b.hIn, b.hOut, b.hReq, b.hBlockReq};
  for (void* p : h)
    if (p) (void)hipHostFree(p);
  if (b.ev) (void)hipEventDestroy(b.ev);
  if (b.s) (void)hipStreamDestroy(b.s);
  b = BatchSlot{};
}

// slots of the workgroup-aligned layout: every request may waste up to kBatchWG - 1
inline size_t aligned_cap(uint32_t cap) { return (size_t)cap + (size_t)kMaxReqPerBatch * kBatchWG; }

hipError_t alloc_slot(BatchSlot& b, uint32_t cap) {
  hipError_t e;
  const size_t acap = aligned_cap(cap), nblk = acap / kBatchWG;
  if ((e = hipStreamCreate(&b.s)) || (e = hipEventCreateWithFlags(&b.ev, hipEventDisableTiming))) return e;
  if ((e = hmalloc(&b.hIn, 3 * (size_t)cap)) || (e = hmalloc(&b.hOut, 6 * (size_t)cap)) ||
      (e = hmalloc(&b.hReq, kMaxReqPerBatch)) || (e = hmalloc(&b.hBlockReq, nblk)))
    return e;
  if ((e = dmalloc(&b.dIn, 3 * (size_t)cap)) || (e = dmalloc(&b.dOut, 6 * (size_t)cap)) || (e = dmalloc(&b.dQ, acap)) ||
      (e = dmalloc(&b.dQn, acap)) || (e = dmalloc(&b.dKey, cap)) || (e = dmalloc(&b.dKey2, cap)) ||
      (e = dmalloc(&b.dIdx, cap)) || (e = dmalloc(&b.dPerm, cap)) || (e = dmalloc(&b.dBlockReq, nblk)) ||
      (e = dmalloc(&b.dReq, kMaxReqPerBatch)))
    return e;
  if ((e = rocprim::radix_sort_pairs(nullptr, b.sortBytes, b.dKey, b.dKey2, b.dIdx, b.dPerm, cap, 0, 40, b.s))) return e;
  return hipMalloc(&b.sortTmp, b.sortBytes + 16);
}

// Takes pending chunks, earliest deadline first, that fit one batch.  Caller holds mu.
void take_requests(mlod_server* S, BatchSlot& b) {
  b.reqs.clear();
  uint32_t slots = 0;
  while (!S->pending.empty() && b.reqs.size() < kMaxReqPerBatch && slots + S->pending.top().m <= S->cap) {
    slots += S->pending.top().m;
    b.reqs.push_back(S->pending.top());
    S->pending.pop();
  }
}

// Stages and launches the requests in b.reqs (worker thread, without the mutex).
int launch_batch(mlod_server* S, BatchSlot& b) {
  // consecutive chunks of one request share a descriptor, so the request's queries are sorted
  // and walked together
  uint32_t off = 0, nReq = 0;
  b.chunkOff.resize(b.reqs.size());
  for (size_t r = 0; r < b.reqs.size(); ++r) {
    const SReq& c = b.reqs[r];
    const Request& q = *c.req;
    b.chunkOff[r] = off;
    std::memcpy(b.hIn + 3 * (size_t)off, q.in + 3 * (size_t)c.off, (size_t)c.m * 3 * sizeof(float));
    const bool extend = r > 0 && b.reqs[r - 1].req == c.req && b.reqs[r - 1].off + b.reqs[r - 1].m == c.off;
    off += c.m;
    if (extend) {
      b.hReq[nReq - 1].count += c.m;
      continue;
    }
    const mlod_tree* t = S->trees[q.tree];
    ReqDesc& d = b.hReq[nReq++];
    d.tv = t->view();
    d.w = make_wcoef(q.kernel);
    d.cutoff = q.cutoff;
    d.stepMax = q.stepMax < 0.f ? t->radius : q.stepMax;
    d.minDepth = q.minDepth;
    d.nIter = q.nIter;
    d.slot0 = b.chunkOff[r];
    d.count = c.m;
    for (int a = 0; a < 3; ++a) {
      const float ext = t->BB[a] - t->bb[a];
      d.org[a] = t->bb[a];
      d.scale[a] = ext > 0.f ? 1024.f / ext : 0.f;
    }
  }
  uint32_t blk = 0;
  for (uint32_t r = 0; r < nReq; ++r) {
    ReqDesc& d = b.hReq[r];
    d.pad0 = blk * kBatchWG;
    const uint32_t nb = (d.count + kBatchWG - 1) / kBatchWG;
    for (uint32_t i = 0; i < nb; ++i) b.hBlockReq[blk + i] = r;
    blk += nb;
  }
  const uint32_t slots = off;
  b.slots = slots;
  int reqBits = 1;
  while ((1u << reqBits) < nReq) ++reqBits;
  hipError_t e = hipMemcpyAsync(b.dIn, b.hIn, (size_t)slots * 3 * sizeof(float), hipMemcpyHostToDevice, b.s);
  if (!e) e = hipMemcpyAsync(b.dReq, b.hReq, nReq * sizeof(ReqDesc), hipMemcpyHostToDevice, b.s);
  if (!e) e = hipMemcpyAsync(b.dBlockReq, b.hBlockReq, blk * sizeof(uint32_t), hipMemcpyHostToDevice, b.s);
  const uint32_t grid = (slots + BS - 1) / BS;
  if (!e) {
    k_batch_keys<<<grid, BS, 0, b.s>>>(b.dIn, slots, b.dReq, nReq, b.dKey, b.dIdx);
    e = hipGetLastError();
  }
  if (!e) e = rocprim::radix_sort_pairs(b.sortTmp, b.sortBytes, b.dKey, b.dKey2, b.dIdx, b.dPerm, slots, 0, 30 + reqBits, b.s);
  if (!e) {
    k_batch_gather<<<grid, BS, 0, b.s>>>(b.dIn, b.dKey2, b.dPerm, slots, b.dReq, b.dQ, b.dQn);
    e = hipGetLastError();
  }
  if (e) return hip_fail(e, "serving batch", __FILE__, __LINE__);
  if (int r = launch_project_batch(b.dReq, b.dBlockReq, b.dQ, b.dQn, 0, blk, S->semantics, S->precision, b.s)) return r;
  k_batch_scatter<<<grid, BS, 0, b.s>>>(b.dQ, b.dQn, b.dKey2, b.dPerm, b.dReq, slots, S->cap, b.dOut);
  e = hipGetLastError();
  if (!e) e = hipMemcpyAsync(b.hOut, b.dOut, (size_t)slots * 3 * sizeof(float), hipMemcpyDeviceToHost, b.s);
  if (!e)
    e = hipMemcpyAsync(b.hOut + 3 * (size_t)S->cap, b.dOut + 3 * (size_t)S->cap, (size_t)slots * 3 * sizeof(float),
                       hipMemcpyDeviceToHost, b.s);
  if (!e) e = hipEventRecord(b.ev, b.s);
  if (e) return hip_fail(e, "serving batch", __FILE__, __LINE__);
  b.launched = clk::now();
  for (const SReq& c : b.reqs)
    if (c.off == 0) c.req->launched = b.launched;
  return MLOD_OK;
}

// Copies a finished batch's results to the callers (worker thread, without the mutex) and
// returns the completion records of the requests it finished.
void complete_batch(mlod_server* S, BatchSlot& b, std::vector<mlod_server_done>& out, std::vector<Request*>& finished) {
  for (size_t r = 0; r < b.reqs.size(); ++r) {
    const SReq& c = b.reqs[r];
    const Request& q = *c.req;
    const uint32_t off = b.chunkOff[r];
    if (q.outP) std::memcpy(q.outP + 3 * (size_t)c.off, b.hOut + 3 * (size_t)off, (size_t)c.m * 3 * sizeof(float));
    if (q.outN)
      std::memcpy(q.outN + 3 * (size_t)c.off, b.hOut + 3 * ((size_t)S->cap + off), (size_t)c.m * 3 * sizeof(float));
  }
  const auto now = clk::now();
  for (const SReq& c : b.reqs) {
    Request& q = *c.req;
    if (--q.chunksLeft) continue;
    mlod_server_done d;
    d.ticket = q.ticket;
    d.latency_ms = std::chrono::duration<double, std::milli>(now - q.submitted).count();
    d.queue_ms = std::chrono::duration<double, std::milli>(q.launched - q.submitted).count();
    d.completed_s = std::chrono::duration<double>(now - S->t0).count();
    out.push_back(d);
    finished.push_back(&q);
  }
}

void set_error(mlod_server* S, int code, const std::string& msg) {
  std::lock_guard<std::mutex> lk(S->mu);
  if (!S->error) {
    S->errmsg = msg;
    S->error = code;
  }
  S->cvDone.notify_all();
}

void worker_main(mlod_server* S) {
  (void)hipSetDevice(S->device);
  std::vector<mlod_server_done> fin;
  std::vector<Request*> finished;
  while (true) {
    bool progressed = false, anyBusy = false;
    // completions: copy out, then publish
    for (BatchSlot& b : S->slots) {
      if (!b.busy) continue;
      const hipError_t q = hipEventQuery(b.ev);
      if (q == hipSuccess) {
        fin.clear();
        finished.clear();
        complete_batch(S, b, fin, finished);
        b.reqs.clear();
        b.busy = false;
        for (Request* q : finished) delete q;
        {
          std::lock_guard<std::mutex> lk(S->mu);
          S->done.insert(S->done.end(), fin.begin(), fin.end());
          S->requests += fin.size();
        }
        S->cvDone.notify_all();
        progressed = true;
      } else if (q != hipErrorNotReady) {
        set_error(S, MLOD_ERR_HIP, std::string("serving batch failed: ") + hipGetErrorString(q));
        return;
      } else {
        anyBusy = true;
      }
    }
    // launches into free slots
    for (BatchSlot& b : S->slots) {
      if (b.busy) continue;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        take_requests(S, b);
      }
      if (b.reqs.empty()) break;
      const int r = launch_batch(S, b);
      if (r) {
        set_error(S, r, g_last_error());
        return;
      }
      b.busy = true;
      anyBusy = true;
      progressed = true;
      std::lock_guard<std::mutex> lk(S->mu);
      S->batches += 1;
      S->queries += b.slots;
    }
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->stop && S->pending.empty() && !anyBusy) return;
    if (progressed) continue;
    if (anyBusy) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    } else {
      S->cv.wait(lk, [S] { return S->stop || !S->pending.empty(); });
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
  if (max_batch_queries < kBatchWG || inflight == 0 || inflight > 8)
    return fail(MLOD_ERR_INVALID, "bad batch capacity or inflight");
  if (chunk_queries > max_batch_queries) return fail(MLOD_ERR_INVALID, "chunk larger than a batch");
  if (precision != MLOD_PREC_FP32 && precision != MLOD_PREC_FP32C)
    return fail(MLOD_ERR_UNSUPPORTED, "serving runs the FP32 and FP32C forms");
  if (semantics != MLOD_SEM_REF_GPU && semantics != MLOD_SEM_REF_CPU) return fail(MLOD_ERR_INVALID, "bad semantics");
  mlod_server* S = new mlod_server();
  S->cap = max_batch_queries;
  S->chunk = chunk_queries ? chunk_queries : std::max<uint32_t>(kBatchWG, max_batch_queries / 4 / kBatchWG * kBatchWG);
  S->semantics = semantics;
  S->precision = precision;
  (void)hipGetDevice(&S->device);
  S->slots.resize(inflight);
  for (BatchSlot& b : S->slots) {
    const hipError_t e = alloc_slot(b, S->cap);
    if (e) {
      for (BatchSlot& c : S->slots) free_slot(c);
      delete S;
      return hip_fail(e, "server buffers", __FILE__, __LINE__);
    }
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
  // requests left behind by a failed worker
  std::vector<Request*> left;
  for (BatchSlot& b : S->slots)
    for (const SReq& c : b.reqs) left.push_back(c.req);
  for (; !S->pending.empty(); S->pending.pop()) left.push_back(S->pending.top().req);
  std::sort(left.begin(), left.end());
  left.erase(std::unique(left.begin(), left.end()), left.end());
  for (Request* q : left) delete q;
  for (BatchSlot& b : S->slots) free_slot(b);
  delete S;
  return MLOD_OK;
}

int mlod_server_add_tree(mlod_server* S, mlod_tree* t, uint32_t* tree_id) {
  if (!S || !t || !tree_id) return fail(MLOD_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *tree_id = (uint32_t)S->trees.size();
  S->trees.push_back(t);
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
  if (k->type < 0 || k->type > 2) return fail(MLOD_ERR_INVALID, "bad kernel type");
  if (k->type == MLOD_KERNEL_GAUSSIAN_MULTIPLE && (k->n_sigmas < 1 || k->n_sigmas > kServeMaxSigmas))
    return fail(MLOD_ERR_UNSUPPORTED, "serving takes 1..8 sigmas");
  Request* r = new Request();
  r->tree = tree_id;
  r->kernel = *k;
  r->cutoff = cutoff;
  r->minDepth = min_depth;
  r->nIter = n_iter;
  r->stepMax = step_max;
  r->in = queries;
  r->outP = out_pts;
  r->outN = out_nmls;
  r->m = m;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (tree_id >= S->trees.size()) {
      delete r;
      return fail(MLOD_ERR_INVALID, "unknown tree id");
    }
    r->ticket = S->nextTicket++;
    r->submitted = clk::now();
    r->launched = r->submitted;
    const double now_ms = std::chrono::duration<double, std::milli>(r->submitted - S->t0).count();
    const double deadline = now_ms + (double)m * std::max<uint32_t>(n_iter, 1u) / kNominalPtsPerMs;
    r->chunksLeft = (m + S->chunk - 1) / S->chunk;
    for (uint32_t off = 0; off < m; off += S->chunk)
      S->pending.push(SReq{r, off, std::min(S->chunk, m - off), deadline, S->nextSeq++});
    *ticket = r->ticket;
  }
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
  out4[0] = S->batches;
  out4[1] = S->requests;
  out4[2] = S->queries;
  out4[3] = S->pending.size();
  return MLOD_OK;
}

}  // extern "C"

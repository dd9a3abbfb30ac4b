// noalloc.hip — mlod_project_points_noalloc: the batched projection of mlod_project_points for
// callers that cannot allocate device memory (code under HIP graph capture, tight serving loops).
//
// mlod_project_points keeps the tree's uploaded queries aside and projects the given points in
// buffers of its own, allocated per call: the Morton-sorted points and their normals, the sort
// permutation, the Morton codes and caller indices, and rocPRIM's sort storage.  Here those
// buffers are carved from one file-scope device array, `scratch`, of kScratchBytes, which holds
// up to MLOD_NOALLOC_MAX_QUERIES points.  The array exists once per GPU for the life of the
// process; its address on each device is resolved with hipGetSymbolAddress on the first call
// on that device (the array's name in host code is only a host-side handle).  Every call on
// one GPU uses that GPU's copy, so calls that overlap on one GPU (two streams, two threads)
// race; calls on one stream are ordered.  The call only launches work on the stream: results
// are in the caller's device arrays once the stream has completed.
#include <mutex>

#include "mlod_internal.h"

namespace mlod {
namespace {

constexpr uint32_t kMaxQ = MLOD_NOALLOC_MAX_QUERIES;
// sorted points and normals (2 x 16 B), permutation, two code arrays and caller indices
// (4 x 4 B) per query, then the sort's temporary storage
constexpr size_t kQBytes = (size_t)kMaxQ * 16;
constexpr size_t kIBytes = (size_t)kMaxQ * 4;
constexpr size_t kSortReserve = (size_t)4 << 20;
constexpr size_t kScratchBytes = 2 * kQBytes + 4 * kIBytes + kSortReserve;  // 16 MiB

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {nullptr};

// The device address of this device's scratch array (resolved once per device).
int scratch_ptr(unsigned char** out) {
  int dev = 0;
  MLOD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(MLOD_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    if (!p) return fail(MLOD_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_ptr[dev];
  return MLOD_OK;
}

}  // namespace
}  // namespace mlod

using namespace mlod;

extern "C" int mlod_project_points_noalloc(mlod_tree* t, const float* d_xyz, uint32_t m, uint32_t n_iter,
                                           float step_max, float* d_out_pts, float* d_out_nmls, void* stream) {
  if (!t) return fail(MLOD_ERR_INVALID, "tree is NULL");
  if (m == 0) return MLOD_OK;
  if (!d_xyz) return fail(MLOD_ERR_INVALID, "points are NULL");
  if (m > kMaxQ)
    return fail(MLOD_ERR_INVALID, "mlod_project_points_noalloc: " + std::to_string(m) + " points exceed the maximum of " +
                                      std::to_string(kMaxQ) + " (MLOD_NOALLOC_MAX_QUERIES)");
  if (!t->kernel_set) return fail(MLOD_ERR_STATE, "no kernel set (mlod_set_kernel)");
  const size_t sortBytes = queries_sort_bytes(m);
  if (sortBytes == 0) return fail(MLOD_ERR_HIP, "query sort: temporary storage size query failed");
  if (sortBytes > kSortReserve) return fail(MLOD_ERR_UNSUPPORTED, "query sort storage exceeds the scratch reserve");
  unsigned char* base = nullptr;
  if (int r = scratch_ptr(&base)) return r;
  if (step_max < 0.f) step_max = t->radius;
  hipStream_t s = (hipStream_t)stream;
  float4* q = reinterpret_cast<float4*>(base);
  float4* qn = reinterpret_cast<float4*>(base + kQBytes);
  uint32_t* perm = reinterpret_cast<uint32_t*>(base + 2 * kQBytes);
  const SortBufs sb{reinterpret_cast<uint32_t*>(base + 2 * kQBytes + kIBytes),
                    reinterpret_cast<uint32_t*>(base + 2 * kQBytes + 2 * kIBytes),
                    reinterpret_cast<uint32_t*>(base + 2 * kQBytes + 3 * kIBytes), base + 2 * kQBytes + 4 * kIBytes,
                    kSortReserve};
  if (int r = queries_order(t, d_xyz, m, q, qn, perm, sb, s)) return r;
  // the uploaded queries are set aside while these points are projected in the scratch buffers;
  // the launches hold their own copies of the pointers, so the state is restored at once
  const QueryState saved = t->qs;
  QueryState tmp{};
  tmp.q = q, tmp.qn = qn, tmp.perm = perm, tmp.m = m, tmp.cap = m;
  t->qs = tmp;
  int r = project_uploaded(t, m, n_iter, step_max, s);
  if (r == MLOD_OK) r = queries_scatter(t, m, d_out_pts, d_out_nmls, s);
  t->qs = saved;
  return r;
}

// queries.hip — device-resident query points (APSS::copyPointsToGPU / project / download,
// sources/apss/GPUUtils.cu:406-411, Fastapss.cu:439-455).
//
// The reference sorts each query batch along a Morton curve on the host before upload
// (Pn::order, sources/apss/Pn.cpp:388-445) so that neighbouring GPU threads traverse the
// same part of the octree.  Here the library does it on the device: a 30-bit Morton code of
// each query in the tree box, a radix sort of (code, caller index), and a gather.  The
// queries then stay in that order, projected in place across calls; downloads scatter back
// to the caller's order through `perm`.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include <vector>

#include "mlod_internal.h"

namespace mlod {
namespace {

constexpr int BS = 256;

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ __launch_bounds__(BS) void k_morton(const float* __restrict__ xyz, uint32_t m, float bx, float by, float bz,
                                               float sx, float sy, float sz, uint32_t* __restrict__ code,
                                               uint32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= m) return;
  const float p[3] = {xyz[3 * (uint64_t)i], xyz[3 * (uint64_t)i + 1], xyz[3 * (uint64_t)i + 2]};
  const float o[3] = {bx, by, bz}, sc[3] = {sx, sy, sz};
  uint32_t c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float u = (p[a] - o[a]) * sc[a];
    u = fminf(fmaxf(u, 0.f), 1023.f);  // NaN -> 0
    c[a] = (uint32_t)u;
  }
  code[i] = spread10(c[0]) | (spread10(c[1]) << 1) | (spread10(c[2]) << 2);
  idx[i] = i;
}

__global__ __launch_bounds__(BS) void k_gather(const float* __restrict__ xyz, const uint32_t* __restrict__ perm, uint32_t m,
                                               float4* __restrict__ q, float4* __restrict__ qn) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= m) return;
  const uint32_t s = perm[i];
  q[i] = make_float4(xyz[3 * (uint64_t)s], xyz[3 * (uint64_t)s + 1], xyz[3 * (uint64_t)s + 2], 0.f);
  qn[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(BS) void k_scatter(const float4* __restrict__ src, const uint32_t* __restrict__ perm, uint32_t m,
                                                uint32_t limit, float* __restrict__ out) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= m) return;
  const uint32_t d = perm[i];
  if (d >= limit) return;
  const float4 v = src[i];
  out[3 * (uint64_t)d] = v.x;
  out[3 * (uint64_t)d + 1] = v.y;
  out[3 * (uint64_t)d + 2] = v.z;
}

// the fit kind of the last step (stored in the normal's w) in the caller's order
__global__ __launch_bounds__(BS) void k_scatter_kind(const float4* __restrict__ qn, const uint32_t* __restrict__ perm,
                                                     uint32_t m, uint32_t limit, uint8_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= m) return;
  const uint32_t d = perm[i];
  if (d < limit) out[d] = (uint8_t)qn[i].w;
}

// inverse of the sort permutation: caller index -> sorted slot
__global__ __launch_bounds__(BS) void k_invperm(const uint32_t* __restrict__ perm, uint32_t m, uint32_t* __restrict__ inv) {
  const uint32_t j = blockIdx.x * BS + threadIdx.x;
  if (j < m) inv[perm[j]] = j;
}

// APSS::reorganizePoints (GPUUtils.cu:413-450): for i = 0, 1, ..., n - 1 in order, the caller's
// queries invalid[i] and endValid - 1 - i exchange positions; normals stay with their index.
// Positions are exchanged by exchanging the two slots' owners (perm), which keeps the Morton
// order of the slots, and the two slots' normals are exchanged back.  One lane runs the list in
// order, so overlapping pairs resolve as in a sequential loop.  Indices were checked on the host.
__global__ void k_reorganize(const uint32_t* __restrict__ invalid, uint32_t n, uint32_t endValid,
                             uint32_t* __restrict__ perm, uint32_t* __restrict__ inv, float4* __restrict__ qn) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t a = invalid[i], b = endValid - 1u - i;
    if (a == b) continue;
    const uint32_t sa = inv[a], sb = inv[b];
    perm[sa] = b;
    perm[sb] = a;
    inv[a] = sb;
    inv[b] = sa;
    const float4 t = qn[sa];
    qn[sa] = qn[sb];
    qn[sb] = t;
  }
}

}  // namespace

void queries_free(mlod_tree* t) {
  QueryState& qs = t->qs;
  if (qs.q) (void)hipFree(qs.q);
  if (qs.qn) (void)hipFree(qs.qn);
  if (qs.perm) (void)hipFree(qs.perm);
  if (qs.stage) (void)hipFree(qs.stage);
  qs = QueryState{};
}

int queries_set(mlod_tree* t, const float* xyz, uint32_t m, bool device_src, hipStream_t s) {
  QueryState& qs = t->qs;
  if (m > qs.cap) {
    queries_free(t);
    if (hipMalloc((void**)&qs.q, (size_t)m * sizeof(float4)) != hipSuccess ||
        hipMalloc((void**)&qs.qn, (size_t)m * sizeof(float4)) != hipSuccess ||
        hipMalloc((void**)&qs.perm, (size_t)m * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&qs.stage, (size_t)m * 3 * sizeof(float)) != hipSuccess) {
      queries_free(t);
      return fail(MLOD_ERR_OOM, "query buffers: hipMalloc failed for " + std::to_string(m) + " queries");
    }
    qs.cap = m;
  }
  qs.m = m;
  if (m == 0) return MLOD_OK;
  const float* src = xyz;
  if (!device_src) {
    MLOD_HIP(hipMemcpyAsync(qs.stage, xyz, (size_t)m * 3 * sizeof(float), hipMemcpyHostToDevice, s));
    src = qs.stage;
  }
  uint32_t *code = nullptr, *code2 = nullptr, *idx = nullptr;
  void* sortTmp = nullptr;
  const size_t sortBytes = queries_sort_bytes(m);
  if (sortBytes == 0) return fail(MLOD_ERR_HIP, "query sort: temporary storage size query failed");
  if (hipMalloc((void**)&code, (size_t)m * 4) != hipSuccess || hipMalloc((void**)&code2, (size_t)m * 4) != hipSuccess ||
      hipMalloc((void**)&idx, (size_t)m * 4) != hipSuccess || hipMalloc(&sortTmp, sortBytes) != hipSuccess) {
    (void)hipFree(code), (void)hipFree(code2), (void)hipFree(idx), (void)hipFree(sortTmp);
    return fail(MLOD_ERR_OOM, "query sort buffers: hipMalloc failed");
  }
  const SortBufs sb{code, code2, idx, sortTmp, sortBytes};
  int r = queries_order(t, src, m, qs.q, qs.qn, qs.perm, sb, s);
  hipError_t e = r == MLOD_OK ? hipStreamSynchronize(s) : hipSuccess;
  (void)hipFree(code), (void)hipFree(code2), (void)hipFree(idx), (void)hipFree(sortTmp);
  if (r != MLOD_OK) return r;
  if (e != hipSuccess) return hip_fail(e, "query upload", __FILE__, __LINE__);
  return MLOD_OK;
}

size_t queries_sort_bytes(uint32_t m) {
  size_t bytes = 0;
  uint32_t* none = nullptr;
  if (rocprim::radix_sort_pairs(nullptr, bytes, none, none, none, none, m, 0, 30) != hipSuccess) return 0;
  return bytes + 16;
}

int queries_order(const mlod_tree* t, const float* d_xyz, uint32_t m, float4* q, float4* qn, uint32_t* perm,
                  const SortBufs& sb, hipStream_t s) {
  if (m == 0) return MLOD_OK;
  float sc[3];
  for (int a = 0; a < 3; ++a) {
    const float ext = t->BB[a] - t->bb[a];
    sc[a] = ext > 0.f ? 1024.f / ext : 0.f;
  }
  const uint32_t nb = (m + BS - 1) / BS;
  k_morton<<<nb, BS, 0, s>>>(d_xyz, m, t->bb[0], t->bb[1], t->bb[2], sc[0], sc[1], sc[2], sb.code, sb.idx);
  MLOD_HIP(hipGetLastError());
  size_t bytes = sb.tmpBytes;
  MLOD_HIP(rocprim::radix_sort_pairs(sb.tmp, bytes, sb.code, sb.code2, sb.idx, perm, m, 0, 30, s));
  k_gather<<<nb, BS, 0, s>>>(d_xyz, perm, m, q, qn);
  MLOD_HIP(hipGetLastError());
  return MLOD_OK;
}

int queries_get(mlod_tree* t, uint32_t n, float* out_pts, float* out_nmls, hipStream_t s) {
  QueryState& qs = t->qs;
  const uint32_t nb = (qs.m + BS - 1) / BS;
  if (out_pts) {
    k_scatter<<<nb, BS, 0, s>>>(qs.q, qs.perm, qs.m, n, qs.stage);
    MLOD_HIP(hipGetLastError());
    MLOD_HIP(hipMemcpyAsync(out_pts, qs.stage, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
  }
  if (out_nmls) {
    k_scatter<<<nb, BS, 0, s>>>(qs.qn, qs.perm, qs.m, n, qs.stage);
    MLOD_HIP(hipGetLastError());
    MLOD_HIP(hipMemcpyAsync(out_nmls, qs.stage, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
  }
  return MLOD_OK;
}

// positions and normals of the first n queries in the caller's order into device arrays
int queries_scatter(mlod_tree* t, uint32_t n, float* d_pts, float* d_nmls, hipStream_t s) {
  QueryState& qs = t->qs;
  const uint32_t nb = (qs.m + BS - 1) / BS;
  if (d_pts && qs.m) {
    k_scatter<<<nb, BS, 0, s>>>(qs.q, qs.perm, qs.m, n, d_pts);
    MLOD_HIP(hipGetLastError());
  }
  if (d_nmls && qs.m) {
    k_scatter<<<nb, BS, 0, s>>>(qs.qn, qs.perm, qs.m, n, d_nmls);
    MLOD_HIP(hipGetLastError());
  }
  return MLOD_OK;
}

}  // namespace mlod

namespace mlod {

int queries_fit_kinds(mlod_tree* t, uint32_t n, uint8_t* out, hipStream_t s) {
  QueryState& qs = t->qs;
  if (n > qs.m) return fail(MLOD_ERR_INVALID, "n exceeds the number of uploaded queries");
  if (n == 0) return MLOD_OK;
  uint8_t* d = nullptr;
  if (hipMalloc((void**)&d, n) != hipSuccess) return fail(MLOD_ERR_OOM, "fit kind buffer");
  k_scatter_kind<<<(qs.m + BS - 1) / BS, BS, 0, s>>>(qs.qn, qs.perm, qs.m, n, d);
  hipError_t e = hipGetLastError();
  if (!e) e = hipMemcpyAsync(out, d, n, hipMemcpyDeviceToHost, s);
  if (!e) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (e) return hip_fail(e, "fit kinds", __FILE__, __LINE__);
  return MLOD_OK;
}

int queries_reorganize(mlod_tree* t, const uint32_t* invalid, uint32_t n, uint32_t end_valid, hipStream_t s) {
  QueryState& qs = t->qs;
  if (end_valid > qs.m) return fail(MLOD_ERR_INVALID, "end_valid exceeds the number of uploaded queries");
  if (n > end_valid) return fail(MLOD_ERR_INVALID, "more invalid points than valid ones");
  for (uint32_t i = 0; i < n; ++i)
    if (invalid[i] >= qs.m) return fail(MLOD_ERR_INVALID, "invalid point index out of range");
  if (n == 0) return MLOD_OK;
  uint32_t *dInv = nullptr, *dList = nullptr;
  if (hipMalloc((void**)&dInv, (size_t)qs.m * 4) != hipSuccess || hipMalloc((void**)&dList, (size_t)n * 4) != hipSuccess) {
    (void)hipFree(dInv);
    return fail(MLOD_ERR_OOM, "reorganize buffers");
  }
  hipError_t e = hipMemcpyAsync(dList, invalid, (size_t)n * 4, hipMemcpyHostToDevice, s);
  if (!e) {
    k_invperm<<<(qs.m + BS - 1) / BS, BS, 0, s>>>(qs.perm, qs.m, dInv);
    e = hipGetLastError();
  }
  if (!e) {
    k_reorganize<<<1, 1, 0, s>>>(dList, n, end_valid, qs.perm, dInv, qs.qn);
    e = hipGetLastError();
  }
  if (!e) e = hipStreamSynchronize(s);
  (void)hipFree(dInv);
  (void)hipFree(dList);
  if (e) return hip_fail(e, "reorganize", __FILE__, __LINE__);
  return MLOD_OK;
}

}  // namespace mlod

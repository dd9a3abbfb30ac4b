// adaptive.hip — the adaptive-octree driver (OctreeGenerator, sources/scene/upsample.cpp) on the GPU.
//
// OctreeGenerator::generate (upsample.cpp:228-251) refines cells of the [-1, 1]^3 octree level by
// level: it projects every cell centre once onto the surface (one step, unbounded step length,
// find_stable_nodesGPU, upsample.cpp:317-346), keeps the cell enclosing each projection at the
// cell's own depth (the root when the projected normal is degenerate), subdivides those cells
// and adds the 26 neighbours of every new child (compute_leaves, upsample.cpp:274-287).  An
// outer loop restarts from the depth-0 neighbours of the cells reached and adapts the final
// depth to the extent of the depth-0 cells used (getRelativeDepth, upsample.cpp:146-161).
//
// Here the cell sets stay on the device as sorted, unique 64-bit keys whose unsigned order is
// the reference's node_comparator (octree.h:147-154: depth, then z, y, x, signed):
//   key = depth << 60 | (z + 2^19) << 40 | (y + 2^19) << 20 | (x + 2^19)
// Every set operation of the reference (sort + unique, children, neighbour expansion, ancestor
// maps) is a kernel plus a rocPRIM radix sort / unique; the projection is the library's own
// projection of the tree (its kernel, parameters, semantics, form and neighbourhood mode).  Only
// the handful of depth-0 cells of the outer loop live on the host.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <vector>

#include "mlod_internal.h"

namespace mlod {
namespace {

constexpr int BS = 256;
constexpr int kBias = 1 << 19;              // coordinates live in [-2^19, 2^19)
constexpr uint64_t kNone = ~(uint64_t)0;    // empty slot of an expansion
constexpr uint32_t kMaxGenDepth = 14;       // depth field: 4 bits, children at max_depth + 1

__host__ __device__ __forceinline__ uint64_t cell_key(int d, int x, int y, int z) {
  return ((uint64_t)d << 60) | ((uint64_t)(uint32_t)(z + kBias) << 40) | ((uint64_t)(uint32_t)(y + kBias) << 20) |
         (uint64_t)(uint32_t)(x + kBias);
}
__host__ __device__ __forceinline__ int key_depth(uint64_t k) { return (int)(k >> 60); }
__host__ __device__ __forceinline__ int key_x(uint64_t k) { return (int)(k & 0xfffffu) - kBias; }
__host__ __device__ __forceinline__ int key_y(uint64_t k) { return (int)((k >> 20) & 0xfffffu) - kBias; }
__host__ __device__ __forceinline__ int key_z(uint64_t k) { return (int)((k >> 40) & 0xfffffu) - kBias; }
__host__ __device__ __forceinline__ bool coord_ok(int v) { return v >= -kBias && v < kBias; }

// Error flags raised by kernels (bit 0: a coordinate left the key range).
enum { kFlagRange = 1 };

// aabb_octree::center over [-1, 1]^3 (octree.h:946-955): min + w p and min + w (p + 1) with
// w = 2 / 2^d, averaged.  Every one of those float operations is exact for |p| < 2^19 and
// d <= 14, so the centre is exactly (2p + 1) 2^-d - 1, which is what is computed.
__device__ __forceinline__ float cell_centre(int p, int d) {
  return __fadd_rn(__fmul_rn((float)(2 * p + 1), __builtin_ldexpf(1.0f, -d)), -1.0f);
}

__global__ __launch_bounds__(BS) void k_centres(const uint64_t* __restrict__ keys, uint32_t n, float* __restrict__ xyz) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const int d = key_depth(k);
  xyz[3 * (size_t)i] = cell_centre(key_x(k), d);
  xyz[3 * (size_t)i + 1] = cell_centre(key_y(k), d);
  xyz[3 * (size_t)i + 2] = cell_centre(key_z(k), d);
}

// aabb_octree::enclosing (octree.h:957-965) over [-1, 1]^3: floor(((p - (-1)) / 2) / 2^-d).
// Returns false for a non-finite or out-of-int coordinate.
__device__ __forceinline__ bool enclose1(float p, int d, int& out) {
  const float t = __fdiv_rn(__fdiv_rn(__fsub_rn(p, -1.0f), 2.0f), __builtin_ldexpf(1.0f, -d));
  const float f = floorf(t);
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) return false;
  out = (int)f;
  return true;
}

// find_stable_nodesGPU's cell assignment (upsample.cpp:336-343) for the projected queries in
// the tree's sorted order: slot j holds the projection of input cell perm[j].  A degenerate
// normal (|n|^2 < 0.5, glm::length2) or a non-finite projected position gives the root.
__global__ __launch_bounds__(BS) void k_enclose(const float4* __restrict__ q, const float4* __restrict__ qn,
                                                const uint32_t* __restrict__ perm, const uint64_t* __restrict__ in,
                                                uint32_t n, uint64_t* __restrict__ out, uint32_t* __restrict__ flags,
                                                unsigned long long* __restrict__ nonfinite) {
  const uint32_t j = blockIdx.x * BS + threadIdx.x;
  if (j >= n) return;
  const int d = key_depth(in[perm[j]]);
  const float4 nm = qn[j];
  const float l2 = __fadd_rn(__fadd_rn(__fmul_rn(nm.x, nm.x), __fmul_rn(nm.y, nm.y)), __fmul_rn(nm.z, nm.z));
  uint64_t key = cell_key(0, 0, 0, 0);
  if (!(l2 < 0.5f)) {
    const float4 p = q[j];
    int x, y, z;
    if (enclose1(p.x, d, x) && enclose1(p.y, d, y) && enclose1(p.z, d, z)) {
      if (coord_ok(x) && coord_ok(y) && coord_ok(z))
        key = cell_key(d, x, y, z);
      else
        atomicOr(flags, (uint32_t)kFlagRange);
    } else {
      atomicAdd(nonfinite, 1ull);
    }
  }
  out[j] = key;
}

// add_children (upsample.cpp:348-358) followed by expand(current, depth + 1)
// (upsample.cpp:409-419): a cell at `depth` contributes the 4 x 4 x 4 block of depth + 1 cells
// made of its children and their neighbours; a shallower cell (the root and its descendants
// re-entered after a failed projection) contributes its 8 children only.  64 slots per cell.
__global__ __launch_bounds__(BS) void k_expand(const uint64_t* __restrict__ sub, uint32_t n, int depth,
                                               uint64_t* __restrict__ out, uint32_t* __restrict__ flags) {
  const uint64_t t = (uint64_t)blockIdx.x * BS + threadIdx.x;
  if (t >= (uint64_t)n * 64) return;
  const uint32_t i = (uint32_t)(t >> 6), s = (uint32_t)(t & 63);
  const uint64_t k = sub[i];
  const int d = key_depth(k), x = key_x(k), y = key_y(k), z = key_z(k);
  int cx, cy, cz;
  if (d == depth) {
    cx = 2 * x - 1 + (int)(s & 3);
    cy = 2 * y - 1 + (int)((s >> 2) & 3);
    cz = 2 * z - 1 + (int)(s >> 4);
  } else if (s < 8) {
    cx = 2 * x + (int)(s & 1);
    cy = 2 * y + (int)((s >> 1) & 1);
    cz = 2 * z + (int)(s >> 2);
  } else {
    out[t] = kNone;
    return;
  }
  if (!coord_ok(cx) || !coord_ok(cy) || !coord_ok(cz)) {
    atomicOr(flags, (uint32_t)kFlagRange);
    out[t] = kNone;
    return;
  }
  out[t] = cell_key(d + 1, cx, cy, cz);
}

// ancestor_at_depth (octree.h:561-573): repeated floor halving, i.e. an arithmetic shift.
// Cells at or above `depth` are kept as they are (onlySetDepthNodes, upsample.cpp:163-173).
__global__ __launch_bounds__(BS) void k_ancestor(const uint64_t* __restrict__ in, uint32_t n, int depth,
                                                 uint64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = in[i];
  const int d = key_depth(k);
  if (d <= depth) {
    out[i] = k;
    return;
  }
  const int sh = d - depth;
  out[i] = cell_key(depth, key_x(k) >> sh, key_y(k) >> sh, key_z(k) >> sh);
}

// generateFromPoints (upsample.cpp:128-144): the max_depth cell enclosing each point of the
// box (aabb_octree::enclosing, octree.h:957-965): floor(((p - min) / diag) / 2^-d).
__global__ __launch_bounds__(BS) void k_enclose_points(const float* __restrict__ xyz, uint32_t n, int d, float mnx,
                                                       float mny, float mnz, float dgx, float dgy, float dgz,
                                                       uint64_t* __restrict__ out, uint32_t* __restrict__ flags,
                                                       unsigned long long* __restrict__ nonfinite) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const float w = __builtin_ldexpf(1.0f, -d);
  const float mn[3] = {mnx, mny, mnz}, dg[3] = {dgx, dgy, dgz};
  int c[3];
  bool fin = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float f = floorf(__fdiv_rn(__fdiv_rn(__fsub_rn(xyz[3 * (size_t)i + a], mn[a]), dg[a]), w));
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) {
      fin = false;
      c[a] = 0;
    } else {
      c[a] = (int)f;
    }
  }
  if (!fin) {
    atomicAdd(nonfinite, 1ull);
    out[i] = kNone;
    return;
  }
  if (!coord_ok(c[0]) || !coord_ok(c[1]) || !coord_ok(c[2])) {
    atomicOr(flags, (uint32_t)kFlagRange);
    out[i] = kNone;
    return;
  }
  out[i] = cell_key(d, c[0], c[1], c[2]);
}

// aabb_octree::center of a general box (octree.h:946-955): w = diag / 2^d, then
// (min + w p + min + w (p + 1)) / 2 in float, without contraction.
__global__ __launch_bounds__(BS) void k_centres_box(const uint64_t* __restrict__ keys, uint32_t n, float mnx, float mny,
                                                    float mnz, float dgx, float dgy, float dgz, float* __restrict__ xyz) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const int d = key_depth(k);
  const int p[3] = {key_x(k), key_y(k), key_z(k)};
  const float mn[3] = {mnx, mny, mnz}, dg[3] = {dgx, dgy, dgz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float w = __fdiv_rn(dg[a], (float)(1 << d));
    const float lo = __fadd_rn(mn[a], __fmul_rn(w, (float)p[a]));
    const float hi = __fadd_rn(mn[a], __fmul_rn(w, (float)(p[a] + 1)));
    xyz[3 * (size_t)i + a] = __fdiv_rn(__fadd_rn(lo, hi), 2.0f);
  }
}

__global__ __launch_bounds__(BS) void k_decode(const uint64_t* __restrict__ keys, uint32_t n, int32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  out[4 * (size_t)i] = key_depth(k);
  out[4 * (size_t)i + 1] = key_x(k);
  out[4 * (size_t)i + 2] = key_y(k);
  out[4 * (size_t)i + 3] = key_z(k);
}

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + BS - 1) / BS); }

// A growable device array of keys.
struct KeyVec {
  uint64_t* p = nullptr;
  size_t n = 0, cap = 0;
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = cap = 0;
  }
  // room for m keys, keeping the first `keep`
  int reserve(size_t m, size_t keep, hipStream_t s) {
    if (m <= cap) return MLOD_OK;
    const size_t nc = std::max(m, cap + cap / 2);
    uint64_t* q = nullptr;
    if (hipMalloc((void**)&q, nc * sizeof(uint64_t)) != hipSuccess)
      return fail(MLOD_ERR_OOM, "adaptive octree: hipMalloc failed for " + std::to_string(nc) + " cells");
    if (keep) MLOD_HIP(hipMemcpyAsync(q, p, keep * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    MLOD_HIP(hipStreamSynchronize(s));
    if (p) (void)hipFree(p);
    p = q;
    cap = nc;
    return MLOD_OK;
  }
};

}  // namespace
}  // namespace mlod

struct mlod_generator {
  mlod_tree* tree = nullptr;
  mlod::KeyVec leaves, total;   // results, sorted and unique
  mlod::KeyVec a, b, c;         // work sets
  void* sortTmp = nullptr;
  size_t sortBytes = 0;
  uint64_t* dCount = nullptr;   // unique count
  uint32_t* dFlags = nullptr;
  unsigned long long* dNonfinite = nullptr;
  float* dXyz = nullptr;        // centres of the leaves (n x 3), valid after mlod_generator_centres
  size_t xyzCap = 0;
  float* dStage = nullptr;      // centres projected by the current level
  size_t stageCap = 0;
  uint32_t maxRel = 0;
  uint64_t projected = 0, calls = 0, nonfinite = 0;
  bool keepTotal = true;
  float box[6] = {-1.f, -1.f, -1.f, 1.f, 1.f, 1.f};  // frame of the leaves' centres
};

namespace mlod {
namespace {

int KeyVec_append(KeyVec& v, const uint64_t* src, size_t m, hipMemcpyKind kind, hipStream_t s) {
  if (!m) return MLOD_OK;
  if (int r = v.reserve(v.n + m, v.n, s)) return r;
  MLOD_HIP(hipMemcpyAsync(v.p + v.n, src, m * sizeof(uint64_t), kind, s));
  v.n += m;
  return MLOD_OK;
}

int ensure_sort_tmp(mlod_generator* g, size_t n, hipStream_t s) {
  size_t need = 0, need2 = 0;
  MLOD_HIP(rocprim::radix_sort_keys(nullptr, need, (uint64_t*)nullptr, (uint64_t*)nullptr, n, 0, 64, s));
  MLOD_HIP(rocprim::unique(nullptr, need2, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint64_t*)nullptr, n,
                           rocprim::equal_to<uint64_t>(), s));
  need = std::max(need, need2);
  if (need <= g->sortBytes) return MLOD_OK;
  MLOD_HIP(hipStreamSynchronize(s));
  if (g->sortTmp) (void)hipFree(g->sortTmp);
  g->sortTmp = nullptr;
  g->sortBytes = 0;
  if (hipMalloc(&g->sortTmp, need + 256) != hipSuccess) return fail(MLOD_ERR_OOM, "adaptive octree: sort buffer");
  g->sortBytes = need;
  return MLOD_OK;
}

// remove_duplicates (upsample.cpp:28-32) of the n keys in `in` (clobbered) into `out`; drops the
// empty-slot key.  Returns the unique count through *m (waits for the stream).
int sort_unique(mlod_generator* g, KeyVec& in, size_t n, KeyVec& out, size_t* m, hipStream_t s) {
  *m = 0;
  if (!n) {
    out.n = 0;
    return MLOD_OK;
  }
  if (n > 0xffffffffu) return fail(MLOD_ERR_UNSUPPORTED, "adaptive octree: more than 2^32 cells in one set");
  if (int r = ensure_sort_tmp(g, n, s)) return r;
  if (int r = out.reserve(n, 0, s)) return r;
  size_t bytes = g->sortBytes;
  MLOD_HIP(rocprim::radix_sort_keys(g->sortTmp, bytes, in.p, out.p, n, 0, 64, s));
  bytes = g->sortBytes;
  MLOD_HIP(rocprim::unique(g->sortTmp, bytes, out.p, in.p, g->dCount, n, rocprim::equal_to<uint64_t>(), s));
  uint64_t cnt = 0, last = 0;
  MLOD_HIP(hipMemcpyAsync(&cnt, g->dCount, sizeof(cnt), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  if (cnt) {
    MLOD_HIP(hipMemcpyAsync(&last, in.p + cnt - 1, sizeof(last), hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
    if (last == kNone) --cnt;
  }
  std::swap(in, out);  // the unique keys are in `in`'s storage: hand it to `out`
  out.n = cnt;
  in.n = 0;
  *m = cnt;
  return MLOD_OK;
}

int check_flags(mlod_generator* g, hipStream_t s) {
  uint32_t f = 0;
  unsigned long long nf = 0;
  MLOD_HIP(hipMemcpyAsync(&f, g->dFlags, sizeof(f), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipMemcpyAsync(&nf, g->dNonfinite, sizeof(nf), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  g->nonfinite = nf;
  if (f & kFlagRange)
    return fail(MLOD_ERR_UNSUPPORTED, "adaptive octree: a cell coordinate left [-2^19, 2^19) (the surface reaches far "
                                      "outside [-1, 1]^3 at this depth)");
  return MLOD_OK;
}

// The tree's own projection of the n centres in g->dStage: one step, step bound FLT_MAX
// (upsample.cpp:331-333).  Results stay in the tree's sorted query arrays.
int project_centres(mlod_generator* g, uint32_t n, hipStream_t s) {
  mlod_tree* t = g->tree;
  if (int r = queries_set(t, g->dStage, n, true, s)) return r;
  int r = (t->mode == MLOD_MODE_KNN || t->mode == MLOD_MODE_BALL) ? launch_neighbors(t, n, 1, s)
                                                                   : launch_project(t, n, 1, FLT_MAX, s);
  g->projected += n;
  g->calls += 1;
  return r;
}

// find_stable_nodesGPU (upsample.cpp:317-346) on the cells in cur -> sorted unique cells in sub
int find_stable(mlod_generator* g, KeyVec& cur, KeyVec& sub, KeyVec& scratch, hipStream_t s) {
  const size_t n = cur.n;
  if (!n) {
    sub.n = 0;
    return MLOD_OK;
  }
  if (n > g->stageCap) {
    if (g->dStage) (void)hipFree(g->dStage);
    g->dStage = nullptr;
    g->stageCap = 0;
    if (hipMalloc((void**)&g->dStage, n * 3 * sizeof(float)) != hipSuccess)
      return fail(MLOD_ERR_OOM, "adaptive octree: centre buffer");
    g->stageCap = n;
  }
  k_centres<<<blocks(n), BS, 0, s>>>(cur.p, (uint32_t)n, g->dStage);
  MLOD_HIP(hipGetLastError());
  if (int r = project_centres(g, (uint32_t)n, s)) return r;
  if (int r = scratch.reserve(n, 0, s)) return r;
  const QueryState& qs = g->tree->qs;
  k_enclose<<<blocks(n), BS, 0, s>>>(qs.q, qs.qn, qs.perm, cur.p, (uint32_t)n, scratch.p, g->dFlags, g->dNonfinite);
  MLOD_HIP(hipGetLastError());
  size_t m = 0;
  if (int r = sort_unique(g, scratch, n, sub, &m, s)) return r;
  return check_flags(g, s);
}

// compute_leaves (upsample.cpp:274-287): cur in, the cells after the last level in cur
int compute_leaves(mlod_generator* g, uint32_t maxDepth, hipStream_t s) {
  KeyVec &cur = g->a, &sub = g->b, &tmp = g->c;
  for (uint32_t depth = 0; depth < maxDepth + 1; ++depth) {
    if (int r = find_stable(g, cur, sub, tmp, s)) return r;
    const size_t ns = sub.n;
    if (int r = tmp.reserve(ns * 64, 0, s)) return r;
    if (ns) {
      k_expand<<<blocks(ns * 64), BS, 0, s>>>(sub.p, (uint32_t)ns, (int)depth, tmp.p, g->dFlags);
      MLOD_HIP(hipGetLastError());
    }
    size_t m = 0;
    if (int r = sort_unique(g, tmp, ns * 64, cur, &m, s)) return r;
    if (int r = check_flags(g, s)) return r;
    if (g->keepTotal)
      if (int r = KeyVec_append(g->total, cur.p, cur.n, hipMemcpyDeviceToDevice, s)) return r;
  }
  return MLOD_OK;
}

// onlySetDepthNodes (upsample.cpp:163-173): ancestors at `depth` of deeper cells, sorted unique
int only_set_depth(mlod_generator* g, KeyVec& v, uint32_t depth, hipStream_t s) {
  if (!v.n) return MLOD_OK;
  KeyVec& tmp = g->c;
  if (int r = tmp.reserve(v.n, 0, s)) return r;
  k_ancestor<<<blocks(v.n), BS, 0, s>>>(v.p, (uint32_t)v.n, (int)depth, tmp.p);
  MLOD_HIP(hipGetLastError());
  size_t m = 0;
  return sort_unique(g, tmp, v.n, v, &m, s);
}

struct HostCell {
  int d, x, y, z;
};

// OctreeGenerator::generate (upsample.cpp:228-251)
int generate(mlod_generator* g, uint32_t maxDepth, hipStream_t s) {
  g->leaves.n = g->total.n = 0;
  g->projected = g->calls = g->nonfinite = 0;
  MLOD_HIP(hipMemsetAsync(g->dFlags, 0, sizeof(uint32_t), s));
  MLOD_HIP(hipMemsetAsync(g->dNonfinite, 0, sizeof(unsigned long long), s));
  std::vector<uint64_t> depth0{cell_key(0, 0, 0, 0)}, old0;
  uint32_t maxRel = maxDepth;
  int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {-INT_MAX, -INT_MAX, -INT_MAX};
  while (!depth0.empty()) {
    if (g->keepTotal)
      if (int r = KeyVec_append(g->total, depth0.data(), depth0.size(), hipMemcpyHostToDevice, s)) return r;
    if (int r = g->a.reserve(depth0.size(), 0, s)) return r;
    MLOD_HIP(hipMemcpyAsync(g->a.p, depth0.data(), depth0.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    g->a.n = depth0.size();
    if (int r = compute_leaves(g, maxRel, s)) return r;
    KeyVec& cur = g->a;
    // depth0_ancestors(current) (upsample.cpp:360-368): a small set, brought to the host
    std::vector<uint64_t> anc;
    if (cur.n) {
      if (int r = g->b.reserve(cur.n, 0, s)) return r;
      k_ancestor<<<blocks(cur.n), BS, 0, s>>>(cur.p, (uint32_t)cur.n, 0, g->b.p);
      MLOD_HIP(hipGetLastError());
      size_t m = 0;
      if (int r = sort_unique(g, g->b, cur.n, g->c, &m, s)) return r;
      anc.resize(m);
      if (m) {
        MLOD_HIP(hipMemcpyAsync(anc.data(), g->c.p, m * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        MLOD_HIP(hipStreamSynchronize(s));
      }
    }
    // getRelativeDepth (upsample.cpp:146-161); min / max persist across iterations
    int rd = 0;
    if (cur.n) {
      for (uint64_t k : anc) {
        const int p[3] = {key_x(k), key_y(k), key_z(k)};
        for (int a = 0; a < 3; ++a) mn[a] = std::min(mn[a], p[a]), mx[a] = std::max(mx[a], p[a]);
      }
      for (int a = 0; a < 3; ++a) rd = std::max(rd, (int)std::ceil(float(std::log2(mx[a] - mn[a] + 1))));
    }
    maxRel = (uint32_t)std::max(0, (int)maxDepth - rd);
    if (int r = KeyVec_append(g->leaves, cur.p, cur.n, hipMemcpyDeviceToDevice, s)) return r;
    old0.insert(old0.end(), depth0.begin(), depth0.end());
    std::sort(old0.begin(), old0.end());
    // expand(depth0) over all depths, minus every depth-0 cell already processed
    std::vector<uint64_t> nxt;
    for (uint64_t k : anc)
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const int x = key_x(k) + dx, y = key_y(k) + dy, z = key_z(k) + dz;
            if (!coord_ok(x) || !coord_ok(y) || !coord_ok(z))
              return fail(MLOD_ERR_UNSUPPORTED, "adaptive octree: depth-0 cell out of range");
            nxt.push_back(cell_key(0, x, y, z));
          }
    std::sort(nxt.begin(), nxt.end());
    nxt.erase(std::unique(nxt.begin(), nxt.end()), nxt.end());
    depth0.clear();
    for (uint64_t k : nxt)
      if (!std::binary_search(old0.begin(), old0.end(), k)) depth0.push_back(k);
  }
  // remove_duplicates(nodes) then onlySetDepthNodes(nodes / total_tree, max relative depth)
  if (int r = only_set_depth(g, g->leaves, maxRel, s)) return r;
  if (g->keepTotal)
    if (int r = only_set_depth(g, g->total, maxRel, s)) return r;
  g->maxRel = maxRel;
  for (int a = 0; a < 3; ++a) g->box[a] = -1.f, g->box[3 + a] = 1.f;
  return check_flags(g, s);
}

int leaf_centres(mlod_generator* g, hipStream_t s) {
  const size_t n = g->leaves.n;
  if (n > g->xyzCap) {
    if (g->dXyz) (void)hipFree(g->dXyz);
    g->dXyz = nullptr;
    g->xyzCap = 0;
    if (hipMalloc((void**)&g->dXyz, std::max<size_t>(n, 1) * 3 * sizeof(float)) != hipSuccess)
      return fail(MLOD_ERR_OOM, "adaptive octree: leaf centre buffer");
    g->xyzCap = n;
  }
  if (n) {
    const float* b = g->box;
    k_centres_box<<<blocks(n), BS, 0, s>>>(g->leaves.p, (uint32_t)n, b[0], b[1], b[2], b[3] - b[0], b[4] - b[1],
                                           b[5] - b[2], g->dXyz);
    MLOD_HIP(hipGetLastError());
  }
  return MLOD_OK;
}

}  // namespace
}  // namespace mlod

using namespace mlod;

extern "C" {

int mlod_generator_create(mlod_tree* t, mlod_generator** out) {
  if (!out) return fail(MLOD_ERR_INVALID, "out is NULL");
  *out = nullptr;
  mlod_generator* g = new mlod_generator();
  g->tree = t;
  if (hipMalloc((void**)&g->dCount, sizeof(uint64_t)) != hipSuccess ||
      hipMalloc((void**)&g->dFlags, sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void**)&g->dNonfinite, sizeof(unsigned long long)) != hipSuccess) {
    mlod_generator_destroy(g);
    return fail(MLOD_ERR_OOM, "generator: hipMalloc failed");
  }
  *out = g;
  return MLOD_OK;
}

int mlod_generator_destroy(mlod_generator* g) {
  if (!g) return MLOD_OK;
  (void)hipDeviceSynchronize();
  g->leaves.release(), g->total.release(), g->a.release(), g->b.release(), g->c.release();
  void* p[] = {g->sortTmp, g->dCount, g->dFlags, g->dNonfinite, g->dXyz, g->dStage};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete g;
  return MLOD_OK;
}

int mlod_generate(mlod_generator* g, uint32_t max_depth, int32_t keep_total, void* stream) {
  if (!g) return fail(MLOD_ERR_INVALID, "generator is NULL");
  if (!g->tree) return fail(MLOD_ERR_STATE, "generator has no tree (created for generate_from_points only)");
  if (max_depth + 1 >= kMaxGenDepth) return fail(MLOD_ERR_INVALID, "max_depth must be <= 12");
  if (!g->tree->kernel_set) return fail(MLOD_ERR_STATE, "no kernel set on the tree (mlod_set_kernel)");
  g->keepTotal = keep_total != 0;
  return generate(g, max_depth, (hipStream_t)stream);
}

int mlod_generate_from_points(mlod_generator* g, const float* xyz, uint32_t n, int32_t device_src, uint32_t max_depth,
                              const float* box6, void* stream) {
  if (!g || (!xyz && n)) return fail(MLOD_ERR_INVALID, "NULL argument");
  if (max_depth >= kMaxGenDepth) return fail(MLOD_ERR_INVALID, "max_depth must be <= 13");
  hipStream_t s = (hipStream_t)stream;
  float b[6] = {-1.f, -1.f, -1.f, 1.f, 1.f, 1.f};
  if (box6) {
    bool empty = false;  // Eigen's isEmpty: some min > max
    for (int a = 0; a < 3; ++a) empty |= box6[a] > box6[3 + a];
    if (!empty) std::memcpy(b, box6, sizeof(b));
  }
  g->leaves.n = g->total.n = 0;
  g->maxRel = max_depth;
  g->projected = g->calls = 0;
  std::memcpy(g->box, b, sizeof(b));
  if (n == 0) return MLOD_OK;  // upsample.cpp:130
  MLOD_HIP(hipMemsetAsync(g->dFlags, 0, sizeof(uint32_t), s));
  MLOD_HIP(hipMemsetAsync(g->dNonfinite, 0, sizeof(unsigned long long), s));
  const float* src = xyz;
  if (!device_src) {
    if (n > g->stageCap) {
      if (g->dStage) (void)hipFree(g->dStage);
      g->dStage = nullptr;
      g->stageCap = 0;
      if (hipMalloc((void**)&g->dStage, (size_t)n * 3 * sizeof(float)) != hipSuccess)
        return fail(MLOD_ERR_OOM, "generator: point buffer");
      g->stageCap = n;
    }
    MLOD_HIP(hipMemcpyAsync(g->dStage, xyz, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
    src = g->dStage;
  }
  if (int r = g->c.reserve(n, 0, s)) return r;
  k_enclose_points<<<blocks(n), BS, 0, s>>>(src, n, (int)max_depth, b[0], b[1], b[2], b[3] - b[0], b[4] - b[1],
                                            b[5] - b[2], g->c.p, g->dFlags, g->dNonfinite);
  MLOD_HIP(hipGetLastError());
  size_t m = 0;
  if (int r = sort_unique(g, g->c, n, g->leaves, &m, s)) return r;
  return check_flags(g, s);
}

int mlod_generator_info(const mlod_generator* g, uint64_t* out6) {
  if (!g || !out6) return fail(MLOD_ERR_INVALID, "NULL argument");
  out6[0] = g->leaves.n;
  out6[1] = g->total.n;
  out6[2] = g->maxRel;
  out6[3] = g->projected;
  out6[4] = g->calls;
  out6[5] = g->nonfinite;
  return MLOD_OK;
}

int mlod_generator_cells(mlod_generator* g, int32_t which, int32_t* out_ids4, void* stream) {
  if (!g || !out_ids4) return fail(MLOD_ERR_INVALID, "NULL argument");
  const KeyVec& v = which ? g->total : g->leaves;
  if (!v.n) return MLOD_OK;
  hipStream_t s = (hipStream_t)stream;
  int32_t* d = nullptr;
  if (hipMalloc((void**)&d, v.n * 4 * sizeof(int32_t)) != hipSuccess) return fail(MLOD_ERR_OOM, "cell download buffer");
  k_decode<<<blocks(v.n), BS, 0, s>>>(v.p, (uint32_t)v.n, d);
  hipError_t e = hipGetLastError();
  if (!e) e = hipMemcpyAsync(out_ids4, d, v.n * 4 * sizeof(int32_t), hipMemcpyDeviceToHost, s);
  if (!e) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (e) return hip_fail(e, "generator cells", __FILE__, __LINE__);
  return MLOD_OK;
}

int mlod_generator_centres(mlod_generator* g, float* out_xyz, const float** d_xyz, void* stream) {
  if (!g) return fail(MLOD_ERR_INVALID, "generator is NULL");
  hipStream_t s = (hipStream_t)stream;
  if (int r = leaf_centres(g, s)) return r;
  if (d_xyz) *d_xyz = g->dXyz;
  if (out_xyz && g->leaves.n) {
    MLOD_HIP(hipMemcpyAsync(out_xyz, g->dXyz, g->leaves.n * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
  }
  return MLOD_OK;
}

}  // extern "C"

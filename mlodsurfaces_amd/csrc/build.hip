// build.hip — octree build, point areas and node moments on the GPU (gfx950).
//
// Restates APSSOctree::build (sources/apss/CPUUtils.cpp:540-569) without recursion:
//   1. root box = float min/max of the points (computeBoundingBox, CPUUtils.cpp:571-579), or the
//      caller's box (updateRootBoundingBox = false, CPUUtils.cpp:547);
//   2. every point descends 12 levels on its own: per level and axis the split is
//      (float)(lo + hi) / 2 (the float sum of CPUUtils.cpp:286-288), the child bit is
//      p > split (CPUUtils.cpp:298-308; exact ties follow tie_policy), the child box
//      takes the split as its new bound (CPUUtils.cpp:319-323), and the child slot is
//      c = xi + 2 yi + 4 zi (CPUUtils.cpp:310).  The 12 slots form a 36-bit path code;
//   3. a stable radix sort of (code, input index) puts the points in leaf order with
//      input order inside a leaf (the stable partition of CPUUtils.cpp:294-311);
//   4. with l(i, j) the number of leading levels two sorted codes share, a node is a leaf iff
//      it holds at most M = maxNumberOfPointsPerLeaf points or sits at maxDepth
//      (CPUUtils.cpp:279-283).  The depth-d node holding point i holds more than M points iff
//      some window of M + 1 consecutive sorted points containing i shares d levels, so i's
//      leaf is at depth L(i) = min(maxDepth, W(i) + 1) with W(i) the largest l(j, j + M) over
//      j = i - M .. i (for M = 1, max(l(i-1,i), l(i,i+1))).  Point i is the first point of the
//      nodes at depths l(i-1,i)+1 .. L(i).  Ranking those "owned" depths per depth gives the
//      breadth-first index (depth, Morton prefix), which is the order of
//      index_nodes_by_breadth_first_search (CPUUtils.cpp:601-619);
//   5. areas = (leaf box diagonal)^2 / #leaf points (estimateAreasRecursive, CPUUtils.cpp:236-252);
//   6. node moments bottom-up, one depth per launch: plain FP32 sums in the reference's
//      order (leaf points in input order, children in slot order, CPUUtils.cpp:358-398),
//      the same in FP64, and node-centred moments for the precision form.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <vector>

#include "mlod_internal.h"

namespace mlod {
namespace {

constexpr int BS = 256;

#define LAUNCH_CHECK()                                                           \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) return hip_fail(e_, "kernel launch", __FILE__, __LINE__); \
  } while (0)

// ----------------------------------------------------------------------------- bbox
__global__ __launch_bounds__(BS) void k_bbox_partial(const float* __restrict__ xyz, uint32_t n, float* __restrict__ part) {
  __shared__ float s[6][BS];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint64_t i = (uint64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BS) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float v = xyz[3 * i + a];
      lo[a] = fminf(lo[a], v);
      hi[a] = fmaxf(hi[a], v);
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s[a][threadIdx.x] = lo[a];
    s[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int st = BS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + st]);
        s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + st]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(BS) void k_bbox_final(const float* __restrict__ part, uint32_t nb, float* __restrict__ out6) {
  __shared__ float s[6][BS];
  float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t b = threadIdx.x; b < nb; b += BS) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      v[a] = fminf(v[a], part[6 * b + a]);
      v[3 + a] = fmaxf(v[3 + a], part[6 * b + 3 + a]);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int st = BS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + st]);
        s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + st]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) out6[threadIdx.x] = s[threadIdx.x][0];
}

// ----------------------------------------------------------------------------- path codes
// One split of one axis.  (double)s * 0.5 is the reference's (bb + BB) / 2.0 with the
// float sum s; the new bound is that value rounded to float (exact for normal floats).
__device__ __forceinline__ uint32_t split_axis(float p, float& lo, float& hi, int policy, int& ties) {
  const float s = __fadd_rn(lo, hi);
  const double mid = (double)s * 0.5;
  const double pd = (double)p;
  uint32_t bit = pd > mid ? 1u : 0u;
  if (pd == mid) {
    ++ties;
    if (policy == 1) bit = 1u;
  }
  const float mf = (float)mid;
  if (bit)
    lo = mf;
  else
    hi = mf;
  return bit;
}

__global__ __launch_bounds__(BS) void k_codes(const float* __restrict__ xyz, uint32_t n, const float* __restrict__ box6,
                                              int maxDepth, int policy, uint64_t* __restrict__ codes,
                                              uint32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  float p[3], lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    p[a] = xyz[3 * (uint64_t)i + a];
    lo[a] = box6[a];
    hi[a] = box6[3 + a];
  }
  uint64_t code = 0;
  int ties = 0;
  for (int l = 0; l < maxDepth; ++l) {
    uint32_t c = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) c |= split_axis(p[a], lo[a], hi[a], policy, ties) << a;
    code = (code << 3) | c;
  }
  codes[i] = code << (3 * (kMaxDepth - maxDepth));
  idx[i] = i;
}

// ----------------------------------------------------------------------------- node ranking
__device__ __forceinline__ int lcp_levels(uint64_t a, uint64_t b, int maxDepth) {
  const uint64_t x = a ^ b;
  if (x == 0) return maxDepth;
  const int l = (__clzll((long long)x) - 28) / 3;
  return l < maxDepth ? l : maxDepth;
}

struct PointSpan {
  int lo;    // first owned depth  (l(i-1,i) + 1)
  int L;     // leaf depth
  int endLo; // first depth whose segment ends at i (l(i,i+1) + 1)
};

// W: the window maxima W(i) for maxNumberOfPointsPerLeaf != 1, or NULL (M = 1).
__device__ __forceinline__ PointSpan point_span(const uint64_t* __restrict__ codes, uint32_t i, uint32_t n, int maxDepth,
                                                const int8_t* __restrict__ W) {
  const uint64_t c = codes[i];
  const int lL = i > 0 ? lcp_levels(codes[i - 1], c, maxDepth) : -1;
  const int lR = i + 1 < n ? lcp_levels(c, codes[i + 1], maxDepth) : -1;
  PointSpan s;
  s.L = min(maxDepth, (W ? (int)W[i] : max(lL, lR)) + 1);
  s.lo = lL + 1;
  s.endLo = lR + 1;
  return s;
}

// A[j] = l(j, j + M), -1 where j + M >= n.
__global__ __launch_bounds__(BS) void k_window_lcp(const uint64_t* __restrict__ codes, uint32_t n, uint32_t M,
                                                   int maxDepth, int8_t* __restrict__ A) {
  const uint32_t j = blockIdx.x * BS + threadIdx.x;
  if (j >= n) return;
  A[j] = (uint64_t)j + M < n ? (int8_t)lcp_levels(codes[j], codes[j + M], maxDepth) : (int8_t)-1;
}

// Running maxima of A forward (P) and backward (S) inside blocks of B = M + 1 entries, one
// thread per block: a window of B entries spans at most two blocks, so its maximum is
// max(S[first], P[last]).
__global__ __launch_bounds__(BS) void k_block_max(const int8_t* __restrict__ A, uint32_t n, uint64_t B,
                                                  int8_t* __restrict__ P, int8_t* __restrict__ S) {
  const uint64_t lo = ((uint64_t)blockIdx.x * BS + threadIdx.x) * B;
  if (lo >= n) return;
  const uint64_t hi = lo + B < n ? lo + B : n;
  int m = -1;
  for (uint64_t j = lo; j < hi; ++j) {
    m = max(m, (int)A[j]);
    P[j] = (int8_t)m;
  }
  m = -1;
  for (uint64_t j = hi; j-- > lo;) {
    m = max(m, (int)A[j]);
    S[j] = (int8_t)m;
  }
}

// Leaf order for maxNumberOfPointsPerLeaf != 1: a leaf above the last level holds points whose
// codes differ below it, and the reference keeps a leaf's points in input order (the stable
// partition of CPUUtils.cpp:294-311).  key = (code cut at the point's leaf depth) << 27 | input
// index, unique, so a sort by it lists the leaves in the same order with their points by index.
__global__ __launch_bounds__(BS) void k_leaf_keys(const uint64_t* __restrict__ codes, const uint32_t* __restrict__ order,
                                                  const int8_t* __restrict__ W, uint32_t n, int maxDepth,
                                                  uint64_t* __restrict__ key, uint32_t* __restrict__ pos) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const int L = min(maxDepth, (int)W[i] + 1);
  const uint64_t keep = L <= 0 ? 0ull : (((1ull << (3 * kMaxDepth)) - 1ull) & ~((1ull << (3 * (kMaxDepth - L))) - 1ull));
  key[i] = ((codes[i] & keep) << 27) | order[i];
  pos[i] = i;
}

// the leaf-ordered codes, input indices and window maxima
__global__ __launch_bounds__(BS) void k_leaf_apply(const uint64_t* __restrict__ key, const uint32_t* __restrict__ pos,
                                                   const int8_t* __restrict__ W, uint32_t n, uint64_t* __restrict__ codes,
                                                   uint32_t* __restrict__ order, int8_t* __restrict__ W2) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  codes[i] = key[i] >> 27;
  order[i] = (uint32_t)(key[i] & ((1u << 27) - 1u));
  W2[i] = W[pos[i]];
}

// W(i) = max of A over j = max(0, i - M) .. i
__global__ __launch_bounds__(BS) void k_window_max(const int8_t* __restrict__ P, const int8_t* __restrict__ S, uint32_t n,
                                                   uint32_t M, int8_t* __restrict__ W) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  W[i] = (uint64_t)i >= M ? (int8_t)max((int)S[i - M], (int)P[i]) : P[i];
}

__global__ __launch_bounds__(BS) void k_count(const uint64_t* __restrict__ codes, uint32_t n, int maxDepth,
                                              const int8_t* __restrict__ W, uint32_t nb, uint32_t* __restrict__ blockCnt) {
  __shared__ uint32_t wt[kMaxDepth + 1][BS / 64];
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  PointSpan sp{1, 0, 1};
  if (i < n) sp = point_span(codes, i, n, maxDepth, W);
#pragma unroll
  for (int d = 0; d <= kMaxDepth; ++d) {
    const bool own = i < n && d <= maxDepth && sp.lo <= d && d <= sp.L;
    const unsigned long long m = __ballot(own);
    if (lane == 0) wt[d][w] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (threadIdx.x <= kMaxDepth) {
    uint32_t s = 0;
    for (int k = 0; k < BS / 64; ++k) s += wt[threadIdx.x][k];
    blockCnt[threadIdx.x * nb + blockIdx.x] = s;
  }
}

// Exclusive scan of the per-block counts of each depth (one workgroup per depth).
__global__ __launch_bounds__(1024) void k_scan_levels(const uint32_t* __restrict__ cnt, uint32_t* __restrict__ off,
                                                      uint32_t nb, uint32_t* __restrict__ levelTot) {
  __shared__ uint32_t s[1024];
  const uint32_t d = blockIdx.x, tid = threadIdx.x;
  const uint32_t* in = cnt + (uint64_t)d * nb;
  uint32_t* out = off + (uint64_t)d * nb;
  const uint32_t chunk = (nb + 1023) / 1024;
  const uint32_t beg = min(nb, tid * chunk), end = min(nb, beg + chunk);
  uint32_t sum = 0;
  for (uint32_t j = beg; j < end; ++j) sum += in[j];
  s[tid] = sum;
  __syncthreads();
  for (uint32_t st = 1; st < 1024; st <<= 1) {
    const uint32_t v = tid >= st ? s[tid - st] : 0u;
    __syncthreads();
    s[tid] += v;
    __syncthreads();
  }
  uint32_t run = s[tid] - sum;
  for (uint32_t j = beg; j < end; ++j) {
    out[j] = run;
    run += in[j];
  }
  if (tid == 1023) levelTot[d] = s[1023];
}

struct LevelOff {
  uint32_t v[kMaxDepth + 2];
};

// Writes every node owned by a point (the point is the first of the node's range) and
// every node whose range ends at the point.
__global__ __launch_bounds__(BS) void k_emit(const uint64_t* __restrict__ codes, const uint32_t* __restrict__ order,
                                             const float* __restrict__ xyz, uint32_t n, int maxDepth,
                                             const int8_t* __restrict__ W, uint32_t nb,
                                             const uint32_t* __restrict__ blockOff, LevelOff lev,
                                             const float* __restrict__ box6, float4* __restrict__ geo,
                                             uint4* __restrict__ hdr, float* __restrict__ box,
                                             uint32_t* __restrict__ nodeLast, uint32_t* __restrict__ pointLeaf,
                                             uint32_t* __restrict__ span,
                                             unsigned long long* __restrict__ tiesOut) {
  __shared__ uint32_t wt[kMaxDepth + 1][BS / 64];
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool valid = i < n;
  PointSpan sp{1, 0, 1};
  if (valid) sp = point_span(codes, i, n, maxDepth, W);
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t rank[kMaxDepth + 1];
  bool own[kMaxDepth + 1];
#pragma unroll
  for (int d = 0; d <= kMaxDepth; ++d) {
    own[d] = valid && d <= maxDepth && sp.lo <= d && d <= sp.L;
    const unsigned long long m = __ballot(own[d]);
    rank[d] = (uint32_t)__popcll(m & lt);
    if (lane == 0) wt[d][w] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (!valid) return;
#pragma unroll
  for (int d = 0; d <= kMaxDepth; ++d) {
    uint32_t o = blockOff[(uint64_t)d * nb + blockIdx.x];
    for (int k = 0; k < w; ++k) o += wt[d][k];
    rank[d] += o;  // exclusive rank of point i among owners of depth d
  }
  const uint32_t src = order[i];
  float p[3], lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    p[a] = xyz[3 * (uint64_t)src + a];
    lo[a] = box6[a];
    hi[a] = box6[3 + a];
  }
  const uint64_t code = codes[i];
  int ties = 0;
#pragma unroll
  for (int d = 0; d <= kMaxDepth; ++d) {
    if (d > sp.L) continue;
    if (own[d]) {
      const uint32_t node = lev.v[d] + rank[d];
      const bool leaf = d == sp.L;
      const int dn = d < kMaxDepth ? d + 1 : kMaxDepth, dp = d > 0 ? d - 1 : 0;
      const uint32_t first = leaf ? i : lev.v[d + 1] + rank[dn];
      uint32_t father = 0xffffffffu;
      if (d > 0) father = lev.v[dp] + rank[dp] - (sp.lo <= d - 1 ? 0u : 1u);
      float c[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) c[a] = __fdiv_rn(__fadd_rn(lo[a], hi[a]), 2.0f);  // BBOX::center, CPUUtils.h:86
      const float dx = __fsub_rn(lo[0], hi[0]), dy = __fsub_rn(lo[1], hi[1]), dz = __fsub_rn(lo[2], hi[2]);
      const float d2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
      // BBOX::radius, CPUUtils.h:84-85.  sqrt in double then one rounding to float is the
      // correctly rounded float sqrt (53 >= 2 * 24 + 2 bits); the GPU's f32 sqrt is not.
      const float r = (float)sqrt((double)d2) * 0.5f;
      geo[node] = make_float4(c[0], c[1], c[2], r);
      hdr[node] = make_uint4(first, ((uint32_t)d << 27) | (leaf ? kLeafBit : 0u), 0u, father);
      span[2 * (uint64_t)node] = i;  // point i starts the subtrees of the nodes it owns
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        box[6 * (uint64_t)node + a] = lo[a];
        box[6 * (uint64_t)node + 3 + a] = hi[a];
      }
    }
    if (d >= sp.endLo) {
      const uint32_t node = lev.v[d] + rank[d] + (own[d] ? 1u : 0u) - 1u;
      const bool leaf = d == sp.L;
      const int dn = d < kMaxDepth ? d + 1 : kMaxDepth;
      const uint32_t last = leaf ? i : lev.v[d + 1] + rank[dn] + (own[dn] ? 1u : 0u) - 1u;
      nodeLast[node] = last;
      span[2 * (uint64_t)node + 1] = i + 1u;  // and ends those of depth >= endLo
    }
    if (d == sp.L) pointLeaf[i] = lev.v[d] + rank[d] + (own[d] ? 1u : 0u) - 1u;  // the leaf holding i
    if (d < sp.L) {  // the split below this node: the path code fixes the child
      const uint32_t cbits = (uint32_t)(code >> (3 * (kMaxDepth - 1 - d))) & 7u;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float s = __fadd_rn(lo[a], hi[a]);
        const float mf = (float)((double)s * 0.5);
        if (p[a] == mf) ++ties;
        if ((cbits >> a) & 1u)
          lo[a] = mf;
        else
          hi[a] = mf;
      }
    }
  }
  if (ties) atomicAdd(tiesOut, (unsigned long long)ties);
}

__global__ __launch_bounds__(BS) void k_finalize(uint32_t nn, uint4* __restrict__ hdr, const uint32_t* __restrict__ nodeLast) {
  const uint32_t k = blockIdx.x * BS + threadIdx.x;
  if (k >= nn) return;
  uint4 h = hdr[k];
  h.y |= (nodeLast[k] - h.x + 1u);
  hdr[k] = h;
}

// Sorted point records and areas (estimateAreasRecursive, CPUUtils.cpp:236-252):
// leafArea = diagonal * diagonal, sampleArea = leafArea / #points; diagonal = 2 r exactly.
__global__ __launch_bounds__(BS) void k_points(const float* __restrict__ xyz, const float* __restrict__ nrmIn,
                                               const uint32_t* __restrict__ order, uint32_t n,
                                               const uint32_t* __restrict__ pointLeaf, const float4* __restrict__ geo,
                                               const uint4* __restrict__ hdr, float4* __restrict__ pts,
                                               float4* __restrict__ nrm, float* __restrict__ areaOrig) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const uint32_t src = order[i];
  const uint32_t leaf = pointLeaf[i];
  const float diag = __fmul_rn(geo[leaf].w, 2.0f);
  const float leafArea = __fmul_rn(diag, diag);
  const float a = (float)((double)leafArea / (double)hdr_count(hdr[leaf].y));  // correctly rounded f32 division
  pts[i] = make_float4(xyz[3 * (uint64_t)src], xyz[3 * (uint64_t)src + 1], xyz[3 * (uint64_t)src + 2], a);
  nrm[i] = make_float4(nrmIn[3 * (uint64_t)src], nrmIn[3 * (uint64_t)src + 1], nrmIn[3 * (uint64_t)src + 2], 0.f);
  areaOrig[src] = a;
}

// ----------------------------------------------------------------------------- moments
struct M9f {
  float a, px, py, pz, nx, ny, nz, pn, pp;
};
struct M9d {
  double a, px, py, pz, nx, ny, nz, pn, pp;
};

// apssNodeStats(p, n, a) (CPUUtils.h:45) in float without contraction, glm dot order.
__device__ __forceinline__ M9f point_stats_f(float4 P, float4 N) {
  const float a = P.w;
  M9f s;
  s.a = a;
  s.px = __fmul_rn(a, P.x);
  s.py = __fmul_rn(a, P.y);
  s.pz = __fmul_rn(a, P.z);
  s.nx = __fmul_rn(a, N.x);
  s.ny = __fmul_rn(a, N.y);
  s.nz = __fmul_rn(a, N.z);
  const float dpn = __fadd_rn(__fadd_rn(__fmul_rn(P.x, N.x), __fmul_rn(P.y, N.y)), __fmul_rn(P.z, N.z));
  const float dpp = __fadd_rn(__fadd_rn(__fmul_rn(P.x, P.x), __fmul_rn(P.y, P.y)), __fmul_rn(P.z, P.z));
  s.pn = __fmul_rn(a, dpn);
  s.pp = __fmul_rn(a, dpp);
  return s;
}
__device__ __forceinline__ void add_f(M9f& s, const M9f& o) {
  s.a = __fadd_rn(s.a, o.a);
  s.nx = __fadd_rn(s.nx, o.nx);
  s.ny = __fadd_rn(s.ny, o.ny);
  s.nz = __fadd_rn(s.nz, o.nz);
  s.px = __fadd_rn(s.px, o.px);
  s.py = __fadd_rn(s.py, o.py);
  s.pz = __fadd_rn(s.pz, o.pz);
  s.pn = __fadd_rn(s.pn, o.pn);
  s.pp = __fadd_rn(s.pp, o.pp);
}
__device__ __forceinline__ M9d point_stats_d(float4 P, float4 N) {
  const double a = P.w, x = P.x, y = P.y, z = P.z, u = N.x, v = N.y, w = N.z;
  M9d s;
  s.a = a;
  s.px = a * x;
  s.py = a * y;
  s.pz = a * z;
  s.nx = a * u;
  s.ny = a * v;
  s.nz = a * w;
  s.pn = a * ((x * u + y * v) + z * w);
  s.pp = a * ((x * x + y * y) + z * z);
  return s;
}
__device__ __forceinline__ void add_d(M9d& s, const M9d& o) {
  s.a += o.a;
  s.px += o.px;
  s.py += o.py;
  s.pz += o.pz;
  s.nx += o.nx;
  s.ny += o.ny;
  s.nz += o.nz;
  s.pn += o.pn;
  s.pp += o.pp;
}

// Kahan-compensated FP32 accumulator.
struct KSum {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    const float y = __fsub_rn(x, c);
    const float t = __fadd_rn(s, y);
    c = __fsub_rn(__fsub_rn(t, s), y);
    s = t;
  }
  __device__ __forceinline__ float val() const { return __fsub_rn(s, c); }
};

__global__ __launch_bounds__(BS) void k_moments(uint32_t beg, uint32_t end, uint4* __restrict__ hdr,
                                                const float4* __restrict__ pts, const float4* __restrict__ nrm,
                                                float4* __restrict__ momP, float4* __restrict__ momC,
                                                double4* __restrict__ momD, double* __restrict__ AD) {
  const uint32_t k = beg + blockIdx.x * BS + threadIdx.x;
  if (k >= end) return;
  uint4 h = hdr[k];
  const uint32_t first = h.x, cnt = hdr_count(h.y);
  M9f sf{0, 0, 0, 0, 0, 0, 0, 0, 0};
  M9d sd{0, 0, 0, 0, 0, 0, 0, 0, 0};
  float pb[3], Cc = 0.f, Vc = 0.f;
  KSum Nk[3];
  if (hdr_leaf(h.y)) {
    for (uint32_t j = first; j < first + cnt; ++j) {
      const float4 P = pts[j], N = nrm[j];
      add_f(sf, point_stats_f(P, N));
      add_d(sd, point_stats_d(P, N));
    }
    // centred form: two passes over the (few) leaf points
    KSum A, S[3];
    for (uint32_t j = first; j < first + cnt; ++j) {
      const float4 P = pts[j], N = nrm[j];
      A.add(P.w);
      S[0].add(P.w * P.x);
      S[1].add(P.w * P.y);
      S[2].add(P.w * P.z);
      Nk[0].add(P.w * N.x);
      Nk[1].add(P.w * N.y);
      Nk[2].add(P.w * N.z);
    }
    const float Av = A.val();
#pragma unroll
    for (int a = 0; a < 3; ++a) pb[a] = cnt == 1 ? (a == 0 ? pts[first].x : a == 1 ? pts[first].y : pts[first].z)
                                                 : S[a].val() / Av;
    KSum Ck, Vk;
    if (cnt > 1) {
      for (uint32_t j = first; j < first + cnt; ++j) {
        const float4 P = pts[j], N = nrm[j];
        const float dx = P.x - pb[0], dy = P.y - pb[1], dz = P.z - pb[2];
        Ck.add(P.w * (dx * N.x + dy * N.y + dz * N.z));
        Vk.add(P.w * (dx * dx + dy * dy + dz * dz));
      }
    }
    Cc = Ck.val();
    Vc = Vk.val();
  } else {
    // plain sums in child slot order (CPUUtils.cpp:379-396)
    for (uint32_t j = first; j < first + cnt; ++j) {
      const float4 m0 = momP[2 * (uint64_t)j], m1 = momP[2 * (uint64_t)j + 1];
      const M9f o{__uint_as_float(hdr[j].z), m0.x, m0.y, m0.z, m1.x, m1.y, m1.z, m0.w, m1.w};
      add_f(sf, o);
      const double4 d0 = momD[2 * (uint64_t)j], d1 = momD[2 * (uint64_t)j + 1];
      const M9d od{AD[j], d0.x, d0.y, d0.z, d1.x, d1.y, d1.z, d0.w, d1.w};
      add_d(sd, od);
    }
    // centred merge: mean, then parallel-axis terms
    const float4 c00 = momC[2 * (uint64_t)first];
    const float A = sf.a;
    KSum D[3];
    for (uint32_t j = first; j < first + cnt; ++j) {
      const float4 c0 = momC[2 * (uint64_t)j], c1 = momC[2 * (uint64_t)j + 1];
      const float Aj = __uint_as_float(hdr[j].z);
      D[0].add(Aj * (c0.x - c00.x));
      D[1].add(Aj * (c0.y - c00.y));
      D[2].add(Aj * (c0.z - c00.z));
      Nk[0].add(c1.x);
      Nk[1].add(c1.y);
      Nk[2].add(c1.z);
    }
    pb[0] = c00.x + D[0].val() / A;
    pb[1] = c00.y + D[1].val() / A;
    pb[2] = c00.z + D[2].val() / A;
    KSum Ck, Vk;
    for (uint32_t j = first; j < first + cnt; ++j) {
      const float4 c0 = momC[2 * (uint64_t)j], c1 = momC[2 * (uint64_t)j + 1];
      const float Aj = __uint_as_float(hdr[j].z);
      const float dx = c0.x - pb[0], dy = c0.y - pb[1], dz = c0.z - pb[2];
      Ck.add(c0.w);
      Ck.add(dx * c1.x + dy * c1.y + dz * c1.z);
      Vk.add(c1.w);
      Vk.add(Aj * (dx * dx + dy * dy + dz * dz));
    }
    Cc = Ck.val();
    Vc = Vk.val();
  }
  h.z = __float_as_uint(sf.a);
  hdr[k] = h;
  momP[2 * (uint64_t)k] = make_float4(sf.px, sf.py, sf.pz, sf.pn);
  momP[2 * (uint64_t)k + 1] = make_float4(sf.nx, sf.ny, sf.nz, sf.pp);
  momC[2 * (uint64_t)k] = make_float4(pb[0], pb[1], pb[2], Cc);
  momC[2 * (uint64_t)k + 1] = make_float4(Nk[0].val(), Nk[1].val(), Nk[2].val(), Vc);
  momD[2 * (uint64_t)k] = make_double4(sd.px, sd.py, sd.pz, sd.pn);
  momD[2 * (uint64_t)k + 1] = make_double4(sd.nx, sd.ny, sd.nz, sd.pp);
  AD[k] = sd.a;
}

template <class T>
int dalloc(T** p, size_t count, uint64_t* bytes) {
  const size_t b = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc((void**)p, b);
  if (e != hipSuccess) return fail(MLOD_ERR_OOM, "hipMalloc of " + std::to_string(b) + " bytes failed");
  if (bytes) *bytes += b;
  return MLOD_OK;
}

struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  int alloc(T** p, size_t count) {
    int r = dalloc(p, count, nullptr);
    if (r == MLOD_OK) ptrs.push_back(*p);
    return r;
  }
};

}  // namespace

int build_tree(mlod_tree* t, const float* d_xyz, const float* d_nrm, hipStream_t s) {
  const uint32_t n = t->n;
  const int maxDepth = (int)t->opts.max_depth;
  Scratch sc;
  int r;
  // 1. bounding box
  const uint32_t nbb = std::min<uint32_t>(1024u, (n + BS - 1) / BS);
  float *part = nullptr, *dbox = nullptr;
  if ((r = sc.alloc(&part, 6 * (size_t)nbb)) || (r = sc.alloc(&dbox, 6))) return r;
  float hbox[6];
  if (t->opts.use_root_box) {
    std::memcpy(hbox, t->opts.root_box, sizeof(hbox));
    MLOD_HIP(hipMemcpyAsync(dbox, hbox, sizeof(hbox), hipMemcpyHostToDevice, s));
    MLOD_HIP(hipStreamSynchronize(s));
  } else {
    k_bbox_partial<<<nbb, BS, 0, s>>>(d_xyz, n, part);
    LAUNCH_CHECK();
    k_bbox_final<<<1, BS, 0, s>>>(part, nbb, dbox);
    LAUNCH_CHECK();
    MLOD_HIP(hipMemcpyAsync(hbox, dbox, sizeof(hbox), hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
  }
  for (int a = 0; a < 3; ++a) {
    t->bb[a] = hbox[a];
    t->BB[a] = hbox[3 + a];
  }
  // 2. path codes, 3. stable sort
  uint64_t *codes = nullptr, *codes2 = nullptr;
  uint32_t *idx = nullptr;
  if ((r = sc.alloc(&codes, n)) || (r = sc.alloc(&codes2, n)) || (r = sc.alloc(&idx, n))) return r;
  if ((r = dalloc(&t->order, n, &t->bytes))) return r;
  const uint32_t nb = (n + BS - 1) / BS;
  k_codes<<<nb, BS, 0, s>>>(d_xyz, n, dbox, maxDepth, t->opts.tie_policy, codes, idx);
  LAUNCH_CHECK();
  size_t tmpBytes = 0;
  MLOD_HIP(rocprim::radix_sort_pairs(nullptr, tmpBytes, codes, codes2, idx, t->order, n, 0, 3 * kMaxDepth, s));
  void* tmp = nullptr;
  if ((r = sc.alloc((char**)&tmp, tmpBytes))) return r;
  MLOD_HIP(rocprim::radix_sort_pairs(tmp, tmpBytes, codes, codes2, idx, t->order, n, 0, 3 * kMaxDepth, s));
  // 4. node ranks per depth
  uint32_t *blockCnt = nullptr, *blockOff = nullptr, *levelTot = nullptr;
  if ((r = sc.alloc(&blockCnt, (size_t)(kMaxDepth + 1) * nb)) || (r = sc.alloc(&blockOff, (size_t)(kMaxDepth + 1) * nb)) ||
      (r = sc.alloc(&levelTot, kMaxDepth + 1)))
    return r;
  // leaf depths for maxNumberOfPointsPerLeaf != 1: window maxima of l(j, j + M)
  int8_t* W = nullptr;
  const uint32_t M = t->opts.max_pts_per_leaf;
  if (M != 1) {
    int8_t *A = nullptr, *P = nullptr, *S = nullptr;
    if ((r = sc.alloc(&A, n)) || (r = sc.alloc(&P, n)) || (r = sc.alloc(&S, n)) || (r = sc.alloc(&W, n))) return r;
    k_window_lcp<<<nb, BS, 0, s>>>(codes2, n, M, maxDepth, A);
    LAUNCH_CHECK();
    const uint64_t B = (uint64_t)M + 1;
    const uint64_t nblk = ((uint64_t)n + B - 1) / B;
    k_block_max<<<(uint32_t)((nblk + BS - 1) / BS), BS, 0, s>>>(A, n, B, P, S);
    LAUNCH_CHECK();
    k_window_max<<<nb, BS, 0, s>>>(P, S, n, M, W);
    LAUNCH_CHECK();
    // re-sort into leaf order (points of a leaf by input index)
    uint64_t *key = nullptr, *key2 = nullptr;
    uint32_t *pos = nullptr, *pos2 = nullptr;
    int8_t* W2 = nullptr;
    if ((r = sc.alloc(&key, n)) || (r = sc.alloc(&key2, n)) || (r = sc.alloc(&pos, n)) || (r = sc.alloc(&pos2, n)) ||
        (r = sc.alloc(&W2, n)))
      return r;
    k_leaf_keys<<<nb, BS, 0, s>>>(codes2, t->order, W, n, maxDepth, key, pos);
    LAUNCH_CHECK();
    size_t tb = 0;
    MLOD_HIP(rocprim::radix_sort_pairs(nullptr, tb, key, key2, pos, pos2, n, 0, 3 * kMaxDepth + 27, s));
    void* tmp2 = nullptr;
    if ((r = sc.alloc((char**)&tmp2, tb))) return r;
    MLOD_HIP(rocprim::radix_sort_pairs(tmp2, tb, key, key2, pos, pos2, n, 0, 3 * kMaxDepth + 27, s));
    k_leaf_apply<<<nb, BS, 0, s>>>(key2, pos2, W, n, codes2, t->order, W2);
    LAUNCH_CHECK();
    W = W2;
  }
  k_count<<<nb, BS, 0, s>>>(codes2, n, maxDepth, W, nb, blockCnt);
  LAUNCH_CHECK();
  k_scan_levels<<<kMaxDepth + 1, 1024, 0, s>>>(blockCnt, blockOff, nb, levelTot);
  LAUNCH_CHECK();
  uint32_t cnt[kMaxDepth + 1];
  MLOD_HIP(hipMemcpyAsync(cnt, levelTot, sizeof(cnt), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  LevelOff lev{};
  uint64_t tot = 0;
  t->depth = 0;
  for (int d = 0; d <= kMaxDepth; ++d) {
    lev.v[d] = (uint32_t)tot;
    t->level_off[d] = (uint32_t)tot;
    tot += cnt[d];
    if (cnt[d]) t->depth = (uint32_t)d;
  }
  lev.v[kMaxDepth + 1] = (uint32_t)tot;
  t->level_off[kMaxDepth + 1] = (uint32_t)tot;
  // kernels address node records with 32-bit byte offsets (32-byte moment pairs)
  if (tot >= (1ull << 27)) return fail(MLOD_ERR_UNSUPPORTED, "octree has 2^27 nodes or more");
  t->nn = (uint32_t)tot;
  const uint32_t nn = t->nn;
  uint32_t *nodeLast = nullptr, *pointLeaf = nullptr;
  unsigned long long* dties = nullptr;
  float* areaOrig = nullptr;
  if ((r = sc.alloc(&nodeLast, nn)) || (r = sc.alloc(&pointLeaf, n)) || (r = sc.alloc(&dties, 1)) ||
      (r = sc.alloc(&areaOrig, n)))
    return r;
  if ((r = dalloc(&t->geo, nn, &t->bytes)) || (r = dalloc(&t->hdr, nn, &t->bytes)) ||
      (r = dalloc(&t->momP, 2 * (size_t)nn, &t->bytes)) || (r = dalloc(&t->momC, 2 * (size_t)nn, &t->bytes)) ||
      (r = dalloc(&t->momD, 2 * (size_t)nn, &t->bytes)) || (r = dalloc(&t->AD, nn, &t->bytes)) ||
      (r = dalloc(&t->box, 6 * (size_t)nn, &t->bytes)) || (r = dalloc(&t->pts, n, &t->bytes)) ||
      (r = dalloc(&t->nrm, n, &t->bytes)) || (r = dalloc(&t->area, n, &t->bytes)) ||
      (r = dalloc(&t->span, nn, &t->bytes)))
    return r;
  MLOD_HIP(hipMemsetAsync(dties, 0, sizeof(unsigned long long), s));
  k_emit<<<nb, BS, 0, s>>>(codes2, t->order, d_xyz, n, maxDepth, W, nb, blockOff, lev, dbox, t->geo, t->hdr, t->box, nodeLast,
                           pointLeaf, reinterpret_cast<uint32_t*>(t->span), dties);
  LAUNCH_CHECK();
  k_finalize<<<(nn + BS - 1) / BS, BS, 0, s>>>(nn, t->hdr, nodeLast);
  LAUNCH_CHECK();
  // 5. sorted points and areas
  k_points<<<nb, BS, 0, s>>>(d_xyz, d_nrm, t->order, n, pointLeaf, t->geo, t->hdr, t->pts, t->nrm, t->area);
  LAUNCH_CHECK();
  // 6. moments bottom-up
  for (int d = (int)t->depth; d >= 0; --d) {
    const uint32_t beg = lev.v[d], end = lev.v[d + 1];
    if (end <= beg) continue;
    k_moments<<<(end - beg + BS - 1) / BS, BS, 0, s>>>(beg, end, t->hdr, t->pts, t->nrm, t->momP, t->momC, t->momD, t->AD);
    LAUNCH_CHECK();
  }
  unsigned long long hties = 0;
  MLOD_HIP(hipMemcpyAsync(&hties, dties, sizeof(hties), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  t->ties = hties;
  return MLOD_OK;
}

int dump_tree(const mlod_tree* t, int32_t* depth, float* bb6, float* center3, float* radius, int32_t* father,
              int32_t* first_child, int32_t* next_brother, int32_t* num_children, uint32_t* leaf_count,
              uint32_t* leaf_idx, float* areas, float* moments9) {
  const uint32_t nn = t->nn, n = t->n;
  std::vector<float4> geo(nn), momP(2 * (size_t)nn);
  std::vector<uint4> hdr(nn);
  std::vector<float> box(6 * (size_t)nn);
  std::vector<uint32_t> order(n);
  MLOD_HIP(hipDeviceSynchronize());
  MLOD_HIP(hipMemcpy(geo.data(), t->geo, nn * sizeof(float4), hipMemcpyDeviceToHost));
  MLOD_HIP(hipMemcpy(hdr.data(), t->hdr, nn * sizeof(uint4), hipMemcpyDeviceToHost));
  MLOD_HIP(hipMemcpy(momP.data(), t->momP, 2 * (size_t)nn * sizeof(float4), hipMemcpyDeviceToHost));
  MLOD_HIP(hipMemcpy(box.data(), t->box, 6 * (size_t)nn * sizeof(float), hipMemcpyDeviceToHost));
  MLOD_HIP(hipMemcpy(order.data(), t->order, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (areas) MLOD_HIP(hipMemcpy(areas, t->area, n * sizeof(float), hipMemcpyDeviceToHost));
  size_t lo = 0;
  for (uint32_t k = 0; k < nn; ++k) {
    const uint4 h = hdr[k];
    const bool leaf = hdr_leaf(h.y);
    const uint32_t c = hdr_count(h.y);
    if (depth) depth[k] = (int32_t)hdr_depth(h.y);
    if (bb6)
      for (int a = 0; a < 6; ++a) bb6[6 * (size_t)k + a] = box[6 * (size_t)k + a];
    if (center3) {
      center3[3 * (size_t)k] = geo[k].x;
      center3[3 * (size_t)k + 1] = geo[k].y;
      center3[3 * (size_t)k + 2] = geo[k].z;
    }
    if (radius) radius[k] = geo[k].w;
    if (father) father[k] = (int32_t)h.w;
    if (first_child) first_child[k] = leaf ? -1 : (int32_t)h.x;
    if (next_brother) next_brother[k] = (k + 1 < nn && k > 0 && hdr[k + 1].w == h.w) ? (int32_t)(k + 1) : -1;
    if (num_children) num_children[k] = leaf ? 0 : (int32_t)c;
    if (leaf_count) leaf_count[k] = leaf ? c : 0u;
    if (leaf && leaf_idx)
      for (uint32_t j = 0; j < c; ++j) leaf_idx[lo++] = order[h.x + j];
    if (moments9) {
      const float4 m0 = momP[2 * (size_t)k], m1 = momP[2 * (size_t)k + 1];
      float A;
      std::memcpy(&A, &h.z, sizeof(float));
      const float v[9] = {A, m0.x, m0.y, m0.z, m1.x, m1.y, m1.z, m0.w, m1.w};
      for (int j = 0; j < 9; ++j) moments9[9 * (size_t)k + j] = v[j];
    }
  }
  return MLOD_OK;
}

}  // namespace mlod

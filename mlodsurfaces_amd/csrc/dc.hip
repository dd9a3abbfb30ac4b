// dc.hip — dual contouring of the adaptive octree, the reference's consumer of the projection
// (scene::computeDualContouring, sources/scene/scene.cpp:1255-1400), on the GPU.
//
// 1. uniformize (sources/scene/upsample.cpp:262-272): the cells (usually the adaptive octree's
//    crossed cells) plus their depth-0 ancestors are re-rooted under one cell of the unit octree
//    (retreiveCorrectRoot / union_nodes, 369-406), graded (grade, 67-93: the parents of the eight
//    diagonal neighbours of every cell, deepest level first), completed upwards (regularize_up)
//    and to full sibling sets (regularize_siblings), optionally cut at max_depth.  Sets are sorted
//    unique 64-bit keys whose order is node_comparator's (depth, z, y, x), one level pass per depth
//    (emit, rocPRIM radix sort, unique).  Cells outside the root after grading, which the traversal
//    from the root never reaches, are dropped.
// 2. Corner field: the distinct corners of the leaves at the finest depth, mapped to the
//    projection frame (new2bbox, scene.cpp:1300-1315), projected one step with the tree's own
//    projection; value (c - p')·n', NaN where |n'|^2 < 0.5 (scene.cpp:1381-1387).  Or values given
//    by the caller.
// 3. Extraction (dual_contour::extract over minimal_edges::traverse, dual_contouring.h:59-87,
//    octree.h:400-537), one thread per leaf edge instead of a recursion: an edge of leaf L is a
//    minimal edge iff none of the four cells around it is finer than L; it is emitted by the first
//    of the four (in the traversal's order) of L's depth, with the sign test on that cell's edge,
//    the traversal's corner order and triangle rule.  Vertices (compute_vertex, 130-176, with the
//    zero normal field the reference passes: the mean of the edge intersections, the cell centre
//    when there is none or it is not finite) are made for every cell of an emitted edge, numbered
//    in node order; triangles are in (leaf, edge) order.  The set of triangles and the vertex of
//    every cell are those of the reference's traversal; only the numbering order differs.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <vector>

#include "mlod_internal.h"

namespace mlod {
namespace {

constexpr int BS = 256;
constexpr int kBias = 1 << 19;
constexpr uint32_t kNoIdx = 0xffffffffu;
constexpr int kMaxDcDepth = 15;  // 4-bit depth field

__host__ __device__ __forceinline__ uint64_t dkey(int d, int x, int y, int z) {
  return ((uint64_t)d << 60) | ((uint64_t)(uint32_t)(z + kBias) << 40) | ((uint64_t)(uint32_t)(y + kBias) << 20) |
         (uint64_t)(uint32_t)(x + kBias);
}
__host__ __device__ __forceinline__ int kd(uint64_t k) { return (int)(k >> 60); }
__host__ __device__ __forceinline__ int kx(uint64_t k) { return (int)(k & 0xfffffu) - kBias; }
__host__ __device__ __forceinline__ int ky(uint64_t k) { return (int)((k >> 20) & 0xfffffu) - kBias; }
__host__ __device__ __forceinline__ int kz(uint64_t k) { return (int)((k >> 40) & 0xfffffu) - kBias; }
// octree_id::parent (octree.h:546-558): floor division by two
__host__ __device__ __forceinline__ int half(int p) { return p >> 1; }
__host__ __device__ __forceinline__ uint64_t ckey(uint32_t x, uint32_t y, uint32_t z) {
  return ((uint64_t)z << 42) | ((uint64_t)y << 21) | (uint64_t)x;
}

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + BS - 1) / BS); }

__device__ __forceinline__ uint32_t find_key(const uint64_t* __restrict__ keys, uint32_t n, uint64_t k) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (keys[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < n && keys[lo] == k) ? lo : kNoIdx;
}

// ids4 (depth, x, y, z) -> keys; flags bit 0 when a cell does not fit the key
__global__ __launch_bounds__(BS) void k_ids_to_keys(const int32_t* __restrict__ ids, uint32_t n, uint64_t* __restrict__ out,
                                                    uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const int d = ids[4 * (size_t)i], x = ids[4 * (size_t)i + 1], y = ids[4 * (size_t)i + 2], z = ids[4 * (size_t)i + 3];
  const bool ok = d >= 0 && d <= kMaxDcDepth && x >= -kBias / 2 && x < kBias / 2 && y >= -kBias / 2 && y < kBias / 2 &&
                  z >= -kBias / 2 && z < kBias / 2;
  if (!ok) atomicOr(flags, 1u);
  out[i] = ok ? dkey(d, x, y, z) : ~(uint64_t)0;
}

// depth0_ancestors (upsample.cpp:360-367) appended after the n cells
__global__ __launch_bounds__(BS) void k_ancestor0(uint64_t* __restrict__ keys, uint32_t n) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const int d = kd(k);
  keys[n + i] = dkey(0, kx(k) >> d, ky(k) >> d, kz(k) >> d);
}

// first index of each depth 0..16 in sorted keys
__global__ void k_depth_bounds(const uint64_t* __restrict__ keys, uint32_t n, uint32_t* __restrict__ out17) {
  const int d = threadIdx.x;
  if (d > 16) return;
  if (d == 16) {
    out17[16] = n;
    return;
  }
  const uint64_t k = (uint64_t)d << 60;
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (keys[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  out17[d] = lo;
}

// union_nodes's shift (upsample.cpp:393-399): pos -= min * 2^depth, depth += delta
__global__ __launch_bounds__(BS) void k_shift(uint64_t* __restrict__ keys, uint32_t n, int mx, int my, int mz, int delta,
                                              uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const int d = kd(k);
  const int nd = d + delta;
  if (nd > kMaxDcDepth) atomicOr(flags, 2u);
  keys[i] = dkey(nd > kMaxDcDepth ? kMaxDcDepth : nd, kx(k) - mx * (1 << d), ky(k) - my * (1 << d), kz(k) - mz * (1 << d));
}

// grade's add_neighbors (upsample.cpp:69-78): parents of the 8 diagonal neighbours of keys[lo..hi)
__global__ __launch_bounds__(BS) void k_grade_level(const uint64_t* __restrict__ keys, uint32_t lo, uint32_t hi,
                                                    uint64_t* __restrict__ out) {
  const uint32_t i = lo + blockIdx.x * BS + threadIdx.x;
  if (i >= hi) return;
  const uint64_t k = keys[i];
  const int d = kd(k), x = kx(k), y = ky(k), z = kz(k);
  uint64_t* o = out + 8 * (size_t)(i - lo);
#pragma unroll
  for (int c = 0; c < 8; ++c)
    o[c] = dkey(d - 1, half(x + ((c & 1) ? 1 : -1)), half(y + ((c & 2) ? 1 : -1)), half(z + ((c & 4) ? 1 : -1)));
}

// regularize_up one level: the parent of each of keys[lo..hi)
__global__ __launch_bounds__(BS) void k_parent_level(const uint64_t* __restrict__ keys, uint32_t lo, uint32_t hi,
                                                     uint64_t* __restrict__ out) {
  const uint32_t i = lo + blockIdx.x * BS + threadIdx.x;
  if (i >= hi) return;
  const uint64_t k = keys[i];
  out[i - lo] = dkey(kd(k) - 1, half(kx(k)), half(ky(k)), half(kz(k)));
}

// regularize_siblings (upsample.cpp:48-65): all children of the parent of each cell (the root
// for depth-0 cells, as the reference's default-constructed slots)
__global__ __launch_bounds__(BS) void k_siblings(const uint64_t* __restrict__ keys, uint32_t n, uint64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const int d = kd(k);
  uint64_t* o = out + 8 * (size_t)i;
  if (d == 0) {
    for (int c = 0; c < 8; ++c) o[c] = dkey(0, 0, 0, 0);
    return;
  }
  const int px = half(kx(k)), py = half(ky(k)), pz = half(kz(k));
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = dkey(d, 2 * px + (c & 1), 2 * py + ((c >> 1) & 1), 2 * pz + ((c >> 2) & 1));
}

// cells inside the root [0, 2^d)^3 (and not deeper than cutDepth)
__global__ __launch_bounds__(BS) void k_keep(const uint64_t* __restrict__ keys, uint32_t n, int cutDepth,
                                             uint8_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const int d = kd(k), w = 1 << d, x = kx(k), y = ky(k), z = kz(k);
  flags[i] = (x >= 0 && x < w && y >= 0 && y < w && z >= 0 && z < w && d <= cutDepth) ? 1 : 0;
}

// leaf flags: no child 0 in the set (regularized: internal cells have all eight)
__global__ __launch_bounds__(BS) void k_leaf_flags(const uint64_t* __restrict__ keys, uint32_t n, uint8_t* __restrict__ leaf) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const int d = kd(k);
  leaf[i] = (d >= kMaxDcDepth || find_key(keys, n, dkey(d + 1, 2 * kx(k), 2 * ky(k), 2 * kz(k))) == kNoIdx) ? 1 : 0;
}

// the 8 corners of each leaf at the finest depth D
__global__ __launch_bounds__(BS) void k_leaf_corners(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ leaves,
                                                     uint32_t nl, int D, uint64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= nl) return;
  const uint64_t k = keys[leaves[i]];
  const int s = D - kd(k);
  const uint32_t x = (uint32_t)kx(k) << s, y = (uint32_t)ky(k) << s, z = (uint32_t)kz(k) << s, w = 1u << s;
#pragma unroll
  for (int c = 0; c < 8; ++c) out[8 * (size_t)i + c] = ckey(x + ((c & 1) ? w : 0), y + ((c & 2) ? w : 0), z + ((c & 4) ? w : 0));
}

// corner positions in the projection frame: new2bbox (c / 2^D) = (2 diag) (c / 2^D) + (2 origin - 1)
__global__ __launch_bounds__(BS) void k_corner_pos(const uint64_t* __restrict__ ck, uint32_t nc, float w, float4 m00,
                                                   float4 m03, float* __restrict__ xyz) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= nc) return;
  const uint64_t k = ck[i];
  const float c[3] = {(float)(uint32_t)(k & 0x1fffffu), (float)(uint32_t)((k >> 21) & 0x1fffffu),
                      (float)(uint32_t)(k >> 42)};
  xyz[3 * (size_t)i] = __fadd_rn(__fmul_rn(m00.x, __fmul_rn(c[0], w)), m03.x);
  xyz[3 * (size_t)i + 1] = __fadd_rn(__fmul_rn(m00.y, __fmul_rn(c[1], w)), m03.y);
  xyz[3 * (size_t)i + 2] = __fadd_rn(__fmul_rn(m00.z, __fmul_rn(c[2], w)), m03.z);
}

// corner value from the projection in the tree's sorted slot j (input corner perm[j]):
// (c - p')·n' in glm's order, NaN where |n'|^2 < 0.5 (scene.cpp:1381-1387)
__global__ __launch_bounds__(BS) void k_corner_values(const float4* __restrict__ q, const float4* __restrict__ qn,
                                                      const uint32_t* __restrict__ perm, uint32_t nc,
                                                      const float* __restrict__ xyz, float* __restrict__ f) {
  const uint32_t j = blockIdx.x * BS + threadIdx.x;
  if (j >= nc) return;
  const uint32_t i = perm[j];
  const float4 p = q[j], n = qn[j];
  const float l2 = __fadd_rn(__fadd_rn(__fmul_rn(n.x, n.x), __fmul_rn(n.y, n.y)), __fmul_rn(n.z, n.z));
  const float dx = __fsub_rn(xyz[3 * (size_t)i], p.x), dy = __fsub_rn(xyz[3 * (size_t)i + 1], p.y),
              dz = __fsub_rn(xyz[3 * (size_t)i + 2], p.z);
  const float v = __fadd_rn(__fadd_rn(__fmul_rn(dx, n.x), __fmul_rn(dy, n.y)), __fmul_rn(dz, n.z));
  f[i] = l2 < 0.5f ? __builtin_nanf("") : v;
}

struct DcView {
  const uint64_t* keys;
  const uint8_t* leaf;
  uint32_t n;
  const uint64_t* ck;  // sorted corner keys
  const float* cf;     // their values
  uint32_t nc;
  int D;
};

__device__ __forceinline__ float corner_value(const DcView& v, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t i = find_key(v.ck, v.nc, ckey(x, y, z));
  return i == kNoIdx ? __builtin_nanf("") : v.cf[i];
}

// The leaf covering cell (d, x, y, z) of the sorted octree: the cell itself or its deepest
// existing ancestor; kNoIdx outside the root.  *finer is set when the cell exists but is not a leaf.
__device__ __forceinline__ uint32_t covering_leaf(const DcView& v, int d, int x, int y, int z, bool* finer) {
  *finer = false;
  const int w = 1 << d;
  if (x < 0 || y < 0 || z < 0 || x >= w || y >= w || z >= w) return kNoIdx;
  for (; d >= 0; --d, x >>= 1, y >>= 1, z >>= 1) {
    const uint32_t i = find_key(v.keys, v.n, dkey(d, x, y, z));
    if (i == kNoIdx) continue;
    if (!v.leaf[i]) {
      *finer = true;
      return kNoIdx;
    }
    return i;
  }
  return kNoIdx;
}

// One thread per (leaf, edge): the minimal edges this leaf emits, two triangle slots each (node
// indices, kNoIdx when absent), and the cells whose vertices the reference creates.
__global__ __launch_bounds__(BS) void k_min_edges(DcView v, const uint32_t* __restrict__ leaves, uint32_t nl,
                                                  uint32_t* __restrict__ tri, uint8_t* __restrict__ used) {
  const uint32_t t = blockIdx.x * BS + threadIdx.x;
  if (t >= nl * 12u) return;
  const uint32_t li = t / 12u, e = t % 12u;
  uint32_t* out = tri + 6 * (size_t)t;
#pragma unroll
  for (int s = 0; s < 6; ++s) out[s] = kNoIdx;
  const uint32_t self = leaves[li];
  const uint64_t k = v.keys[self];
  const int d = kd(k), p[3] = {kx(k), ky(k), kz(k)};
  // edge e: direction dir, at the cell's (j, k) side in the two other axes (a, b):
  // X: (a, b) = (y, z); Y: (x, z); Z: (x, y) — the traversal's quadrant order n1..n4 is
  // (a-, b-), (a+, b-), (a-, b+), (a+, b+)
  const int dir = (int)(e >> 2), j = (int)(e & 1), kk = (int)((e >> 1) & 1);
  const int A = dir == 0 ? 1 : 0, B = dir == 2 ? 1 : 2;
  // the cell is on the a- side of the edge when the edge is at its high a face
  const int myQ = (j ? 0 : 1) + 2 * (kk ? 0 : 1);
  uint32_t nodes[4];
  int depth[4];
  for (int q = 0; q < 4; ++q) {
    int c[3] = {p[0], p[1], p[2]};
    const int qa = q & 1, qb = q >> 1;
    c[A] = p[A] + j - 1 + qa;  // edge coordinate p[A] + j; the a- cell starts one below it
    c[B] = p[B] + kk - 1 + qb;
    if (q == myQ) {
      nodes[q] = self;
      depth[q] = d;
      continue;
    }
    bool finer = false;
    const uint32_t i = covering_leaf(v, d, c[0], c[1], c[2], &finer);
    if (i == kNoIdx) return;  // a finer cell owns the edge, or the edge is on the root's boundary
    nodes[q] = i;
    depth[q] = kd(v.keys[i]);
  }
  // selected: the first of the deepest (dual_contouring.h:201-206); this cell is one of the
  // deepest, and emits only if it is the first of them
  int sel = 0;
  for (int q = 1; q < 4; ++q)
    if (depth[q] > depth[sel]) sel = q;
  if (sel != myQ) return;
  // the selected cell's edge ends, in corner_indices' order (dual_contouring.h:208-213): low end
  // first along X and Z, high end first along Y
  const int D = v.D, sh = D - d;
  uint32_t c0[3] = {(uint32_t)p[0], (uint32_t)p[1], (uint32_t)p[2]};
  c0[A] += j, c0[B] += kk;
  uint32_t c1[3] = {c0[0], c0[1], c0[2]};
  if (dir == 1)
    c0[dir] += 1;
  else
    c1[dir] += 1;
  const float f0 = corner_value(v, c0[0] << sh, c0[1] << sh, c0[2] << sh);
  const float f1 = corner_value(v, c1[0] << sh, c1[1] << sh, c1[2] << sh);
  if (__builtin_isnan(f0) || __builtin_isnan(f1)) return;
  const bool m0 = f0 < 0.f, m1 = f1 < 0.f;
  if (m0 == m1) return;
  for (int q = 0; q < 4; ++q) used[nodes[q]] = 1;  // add_triangle makes all three vertices first
  // add_triangle(0, 1, 3) and (0, 3, 2) with the flip (dual_contouring.h:70-84)
  const uint32_t ta[3] = {nodes[0], m0 ? nodes[3] : nodes[1], m0 ? nodes[1] : nodes[3]};
  const uint32_t tb[3] = {nodes[0], m0 ? nodes[2] : nodes[3], m0 ? nodes[3] : nodes[2]};
  if (ta[0] != ta[1] && ta[1] != ta[2] && ta[2] != ta[0]) out[0] = ta[0], out[1] = ta[1], out[2] = ta[2];
  if (tb[0] != tb[1] && tb[1] != tb[2] && tb[2] != tb[0]) out[3] = tb[0], out[4] = tb[1], out[5] = tb[2];
}

// compute_vertex (dual_contouring.h:130-176) of each used cell, in its unit frame; vertex ids
// from the exclusive scan of `used`
__global__ __launch_bounds__(BS) void k_vertices(DcView v, const uint8_t* __restrict__ used,
                                                 const uint32_t* __restrict__ vid, float* __restrict__ unitXyz,
                                                 uint32_t* __restrict__ vnode) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= v.n || !used[i]) return;
  const uint64_t k = v.keys[i];
  const int d = kd(k), sh = v.D - d;
  const int p[3] = {kx(k), ky(k), kz(k)};
  const float w = 1.0f / (float)(1 << v.D);
  const uint8_t E[12][2] = {{0, 1}, {0, 2}, {1, 3}, {2, 3}, {4, 5}, {4, 6}, {5, 7}, {6, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  float sum[3] = {0.f, 0.f, 0.f};
  uint32_t count = 0;
  for (int e = 0; e < 12; ++e) {
    uint32_t u[3], q[3];
    for (int a = 0; a < 3; ++a) {
      u[a] = (uint32_t)(p[a] + ((E[e][0] >> a) & 1)) << sh;
      q[a] = (uint32_t)(p[a] + ((E[e][1] >> a) & 1)) << sh;
    }
    const float fu = corner_value(v, u[0], u[1], u[2]), fv = corner_value(v, q[0], q[1], q[2]);
    if ((fu < 0.f) == (fv < 0.f)) continue;
    const float x = __fdiv_rn(-fu, __fsub_rn(fv, fu));
    const float omx = __fsub_rn(1.0f, x);
    for (int a = 0; a < 3; ++a)
      sum[a] = __fadd_rn(sum[a], __fadd_rn(__fmul_rn(x, __fmul_rn((float)q[a], w)), __fmul_rn(omx, __fmul_rn((float)u[a], w))));
    ++count;
  }
  const float width = 1.0f / (float)(1 << d);
  float r[3];
  bool nan = count == 0;
  for (int a = 0; a < 3; ++a) {
    r[a] = __fdiv_rn(sum[a], (float)count);
    nan = nan || __builtin_isnan(r[a]);
  }
  if (nan)
    for (int a = 0; a < 3; ++a) r[a] = __fmul_rn(__fadd_rn((float)p[a], 0.5f), width);
  const uint32_t o = vid[i];
  for (int a = 0; a < 3; ++a) unitXyz[3 * (size_t)o + a] = __fadd_rn(r[a], 0.0f);
  vnode[o] = i;
}

// triangles from node indices to vertex ids
__global__ __launch_bounds__(BS) void k_tri_ids(uint32_t* __restrict__ tri, uint32_t nt, const uint32_t* __restrict__ vid) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= 3 * nt) return;
  tri[i] = vid[tri[i]];
}

// vertices from the unit frame to the projection frame
__global__ __launch_bounds__(BS) void k_to_frame(const float* __restrict__ u, uint32_t nv, float4 m00, float4 m03,
                                                 float* __restrict__ out) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= nv) return;
  out[3 * (size_t)i] = __fadd_rn(__fmul_rn(m00.x, u[3 * (size_t)i]), m03.x);
  out[3 * (size_t)i + 1] = __fadd_rn(__fmul_rn(m00.y, u[3 * (size_t)i + 1]), m03.y);
  out[3 * (size_t)i + 2] = __fadd_rn(__fmul_rn(m00.z, u[3 * (size_t)i + 2]), m03.z);
}

// ----------------------------------------------------------------------------- mesh passes
// mesh::smoothMesh and mesh::subdivideMesh (sources/scene/upsample.h:55-133), in the reference's
// float operations and orders: a stable sort by vertex (smoothing) or by edge (subdivision) keeps
// the triangle order in which the sequential code accumulates and numbers, so the results are
// those of the reference bit for bit.

// encounter e = 3 t + k of triangle t: smoothing key = its k-th vertex; subdivision key = its
// k-th edge (a, b), (c, b), (c, a) as (min, max)
__global__ __launch_bounds__(BS) void k_mesh_keys(const uint32_t* __restrict__ tri, uint32_t nt, int edges,
                                                  uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint32_t e = blockIdx.x * BS + threadIdx.x;
  if (e >= 3 * nt) return;
  const uint32_t t = e / 3, k = e % 3;
  const uint32_t a = tri[3 * (size_t)t], b = tri[3 * (size_t)t + 1], c = tri[3 * (size_t)t + 2];
  if (edges) {
    const uint32_t u = k == 0 ? a : c, v = k == 0 ? b : (k == 1 ? b : a);
    key[e] = ((uint64_t)min(u, v) << 32) | (uint64_t)max(u, v);
  } else {
    key[e] = k == 0 ? a : (k == 1 ? b : c);
  }
  val[e] = e;
}

// smoothMesh: newPos[v] += (the other two vertices of each of v's triangles, in triangle order),
// valence 2 per triangle, then newPos / valence (NaN for an unreferenced vertex, as 0 / 0)
__global__ __launch_bounds__(BS) void k_smooth(const float* __restrict__ P, uint32_t nv, const uint32_t* __restrict__ tri,
                                               const uint64_t* __restrict__ skey, const uint32_t* __restrict__ senc,
                                               uint32_t ne, float* __restrict__ out) {
  const uint32_t v = blockIdx.x * BS + threadIdx.x;
  if (v >= nv) return;
  uint32_t lo = 0, hi = ne;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (skey[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  float s[3] = {0.f, 0.f, 0.f}, val = 0.f;
  for (uint32_t i = lo; i < ne && skey[i] == v; ++i) {
    const uint32_t e = senc[i], t = e / 3, k = e % 3;
    const uint32_t a = tri[3 * (size_t)t], b = tri[3 * (size_t)t + 1], c = tri[3 * (size_t)t + 2];
    // a: P[b] + P[c]; b: P[a] + P[c]; c: P[b] + P[a]
    const uint32_t x = k == 0 ? b : (k == 1 ? a : b), y = k == 0 ? c : (k == 1 ? c : a);
    for (int q = 0; q < 3; ++q) s[q] = __fadd_rn(s[q], __fadd_rn(P[3 * (size_t)x + q], P[3 * (size_t)y + q]));
    val = __fadd_rn(val, 2.0f);
  }
  for (int q = 0; q < 3; ++q) out[3 * (size_t)v + q] = __fdiv_rn(s[q], val);
}

// subdivideMesh, step 1: head of each run of equal edges in the sorted encounters; the first
// encounter of each edge marked in encounter order
__global__ __launch_bounds__(BS) void k_edge_heads(const uint64_t* __restrict__ skey, const uint32_t* __restrict__ senc,
                                                   uint32_t ne, uint32_t* __restrict__ headIdx,
                                                   uint8_t* __restrict__ isFirst) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= ne) return;
  const bool head = i == 0 || skey[i] != skey[i - 1];
  headIdx[i] = head ? i : 0;
  isFirst[senc[i]] = head ? 1 : 0;
}

// step 2: each encounter's new vertex (nv + rank of its edge's first encounter), and the midpoint
// (P[u] + P[v]) / 2 of each edge at its head
__global__ __launch_bounds__(BS) void k_edge_verts(const uint64_t* __restrict__ skey, const uint32_t* __restrict__ senc,
                                                   const uint32_t* __restrict__ headScan, const uint32_t* __restrict__ rank,
                                                   uint32_t ne, uint32_t nv, const float* __restrict__ P,
                                                   uint32_t* __restrict__ newV, float* __restrict__ outP) {
  const uint32_t i = blockIdx.x * BS + threadIdx.x;
  if (i >= ne) return;
  const uint32_t h = headScan[i];
  const uint32_t id = nv + rank[senc[h]];
  newV[senc[i]] = id;
  if (h == i) {
    const uint32_t u = (uint32_t)(skey[i] >> 32), v = (uint32_t)skey[i];
    for (int q = 0; q < 3; ++q)
      outP[3 * (size_t)id + q] = __fdiv_rn(__fadd_rn(P[3 * (size_t)u + q], P[3 * (size_t)v + q]), 2.0f);
  }
}

// step 3: triangle t becomes (a, ab, ca); (ab, b, bc), (ab, bc, ca), (ca, bc, c) are appended
__global__ __launch_bounds__(BS) void k_subdiv_tris(const uint32_t* __restrict__ tri, uint32_t nt,
                                                    const uint32_t* __restrict__ newV, uint32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * BS + threadIdx.x;
  if (t >= nt) return;
  const uint32_t a = tri[3 * (size_t)t], b = tri[3 * (size_t)t + 1], c = tri[3 * (size_t)t + 2];
  const uint32_t ab = newV[3 * t], bc = newV[3 * t + 1], ca = newV[3 * t + 2];
  uint32_t* o = out + 3 * (size_t)t;
  o[0] = a, o[1] = ab, o[2] = ca;
  uint32_t* q = out + 3 * ((size_t)nt + 3 * (size_t)t);
  q[0] = ab, q[1] = b, q[2] = bc;
  q[3] = ab, q[4] = bc, q[5] = ca;
  q[6] = ca, q[7] = bc, q[8] = c;
}

template <class T>
struct DVec {
  T* p = nullptr;
  size_t n = 0, cap = 0;
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = cap = 0;
  }
  int reserve(size_t m, size_t keep, hipStream_t s) {
    if (m <= cap) return MLOD_OK;
    const size_t nc = std::max(m, cap + cap / 2);
    T* q = nullptr;
    if (hipMalloc((void**)&q, std::max<size_t>(nc, 1) * sizeof(T)) != hipSuccess)
      return fail(MLOD_ERR_OOM, "dual contouring: hipMalloc failed for " + std::to_string(nc) + " elements");
    if (keep) MLOD_HIP(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
    MLOD_HIP(hipStreamSynchronize(s));
    if (p) (void)hipFree(p);
    p = q;
    cap = nc;
    return MLOD_OK;
  }
};

}  // namespace
}  // namespace mlod

struct mlod_dc {
  mlod::DVec<uint64_t> keys, work, ck;
  mlod::DVec<uint8_t> leaf, used, flags8;
  mlod::DVec<uint32_t> leaves, vid, tri, vnode, u32;
  mlod::DVec<float> cf, cpos, verts, vframe, rv, rn;
  mlod::DVec<uint32_t> rt;
  uint32_t nRV = 0, nRT = 0;
  bool refined = false;
  void* tmp = nullptr;
  size_t tmpBytes = 0;
  uint32_t* dCount = nullptr;
  uint64_t* dCount64 = nullptr;
  uint32_t* dFlags = nullptr;
  uint32_t* dBounds = nullptr;
  int D = 0;
  float diag = 1.f, origin[3] = {0.f, 0.f, 0.f};
  uint32_t nLeaves = 0, nVerts = 0, nTris = 0;
  bool haveValues = false, extracted = false;
};

namespace mlod {
namespace {

int ensure_tmp(mlod_dc* dc, size_t need) {
  if (need <= dc->tmpBytes) return MLOD_OK;
  (void)hipDeviceSynchronize();
  if (dc->tmp) (void)hipFree(dc->tmp);
  dc->tmp = nullptr;
  dc->tmpBytes = 0;
  if (hipMalloc(&dc->tmp, need + 256) != hipSuccess) return fail(MLOD_ERR_OOM, "dual contouring: temporary storage");
  dc->tmpBytes = need;
  return MLOD_OK;
}

// keys.p[0..n) (any order, clobbered) -> sorted unique keys in dc->keys; returns the count
int sort_unique(mlod_dc* dc, mlod::DVec<uint64_t>& in, size_t n, size_t* m, hipStream_t s) {
  *m = 0;
  if (n == 0) {
    in.n = 0;
    return MLOD_OK;
  }
  if (n > 0xffffffffu) return fail(MLOD_ERR_UNSUPPORTED, "dual contouring: more than 2^32 cells");
  size_t a = 0, b = 0;
  MLOD_HIP(rocprim::radix_sort_keys(nullptr, a, (uint64_t*)nullptr, (uint64_t*)nullptr, n, 0, 64, s));
  MLOD_HIP(rocprim::unique(nullptr, b, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint64_t*)nullptr, n,
                           rocprim::equal_to<uint64_t>(), s));
  if (int r = ensure_tmp(dc, std::max(a, b))) return r;
  if (int r = dc->work.reserve(n, 0, s)) return r;
  size_t bytes = dc->tmpBytes;
  MLOD_HIP(rocprim::radix_sort_keys(dc->tmp, bytes, in.p, dc->work.p, n, 0, 64, s));
  bytes = dc->tmpBytes;
  MLOD_HIP(rocprim::unique(dc->tmp, bytes, dc->work.p, in.p, dc->dCount64, n, rocprim::equal_to<uint64_t>(), s));
  uint64_t cnt = 0;
  MLOD_HIP(hipMemcpyAsync(&cnt, dc->dCount64, sizeof(cnt), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  in.n = cnt;
  *m = cnt;
  return MLOD_OK;
}

int depth_bounds(mlod_dc* dc, uint32_t b[17], hipStream_t s) {
  k_depth_bounds<<<1, 64, 0, s>>>(dc->keys.p, (uint32_t)dc->keys.n, dc->dBounds);
  MLOD_HIP(hipGetLastError());
  MLOD_HIP(hipMemcpyAsync(b, dc->dBounds, 17 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  return MLOD_OK;
}

// appends f(keys at depth d) for every depth from deepest to 1 (each level sees what the deeper
// ones added), keeping dc->keys sorted and unique
template <class Emit>
int level_pass(mlod_dc* dc, uint32_t perNode, Emit emit, hipStream_t s) {
  for (int d = kMaxDcDepth; d >= 1; --d) {
    uint32_t b[17];
    if (int r = depth_bounds(dc, b, s)) return r;
    const uint32_t lo = b[d], hi = b[d + 1];
    if (hi <= lo) continue;
    const size_t n = dc->keys.n, add = (size_t)(hi - lo) * perNode;
    if (int r = dc->keys.reserve(n + add, n, s)) return r;
    emit(dc->keys.p, lo, hi, dc->keys.p + n, s);
    MLOD_HIP(hipGetLastError());
    size_t m = 0;
    if (int r = sort_unique(dc, dc->keys, n + add, &m, s)) return r;
  }
  return MLOD_OK;
}

// keeps the keys whose flag is set (order kept)
int compact_keys(mlod_dc* dc, hipStream_t s) {
  const size_t n = dc->keys.n;
  if (!n) return MLOD_OK;
  size_t need = 0;
  MLOD_HIP(rocprim::select(nullptr, need, (uint64_t*)nullptr, (uint8_t*)nullptr, (uint64_t*)nullptr, dc->dCount, n, s));
  if (int r = ensure_tmp(dc, need)) return r;
  if (int r = dc->work.reserve(n, 0, s)) return r;
  size_t bytes = dc->tmpBytes;
  MLOD_HIP(rocprim::select(dc->tmp, bytes, dc->keys.p, dc->flags8.p, dc->work.p, dc->dCount, n, s));
  uint32_t cnt = 0;
  MLOD_HIP(hipMemcpyAsync(&cnt, dc->dCount, sizeof(cnt), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  std::swap(dc->keys, dc->work);
  dc->keys.n = cnt;
  return MLOD_OK;
}

int check_flags(mlod_dc* dc, hipStream_t s) {
  uint32_t f = 0;
  MLOD_HIP(hipMemcpyAsync(&f, dc->dFlags, sizeof(f), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  if (f & 1) return fail(MLOD_ERR_INVALID, "dual contouring: a cell's depth or coordinates are out of range");
  if (f & 2) return fail(MLOD_ERR_UNSUPPORTED, "dual contouring: re-rooting deepens the octree beyond 15 levels");
  return MLOD_OK;
}

void frame_matrix(const mlod_dc* dc, float4& m00, float4& m03) {
  // new2bbox = unit2bbox * sortedOctreeToUnit (scene.cpp:1262-1315), unit2bbox for [-1, 1]^3
  const float two = 2.0f;
  m00 = make_float4(two * dc->diag, two * dc->diag, two * dc->diag, 0.f);
  m03 = make_float4(two * dc->origin[0] + -1.0f, two * dc->origin[1] + -1.0f, two * dc->origin[2] + -1.0f, 0.f);
}

// uniformize(maxDepth, cells, ..., cut) or, with fromPoints, uniformizeFromPoints(cells)
// (upsample.cpp:253-259: no depth-0 ancestors, no re-rooting, no cut)
int build(mlod_dc* dc, const int32_t* ids4, uint32_t n, uint32_t maxDepth, bool cut, bool fromPoints, hipStream_t s) {
  dc->extracted = dc->haveValues = dc->refined = false;
  dc->nLeaves = dc->nVerts = dc->nTris = 0;
  MLOD_HIP(hipMemsetAsync(dc->dFlags, 0, sizeof(uint32_t), s));
  if (int r = dc->keys.reserve(2 * (size_t)n + 1, 0, s)) return r;
  int32_t* dIds = nullptr;
  if (hipMalloc((void**)&dIds, std::max<size_t>(n, 1) * 4 * sizeof(int32_t)) != hipSuccess)
    return fail(MLOD_ERR_OOM, "dual contouring: cells");
  hipError_t e = hipMemcpyAsync(dIds, ids4, (size_t)n * 4 * sizeof(int32_t), hipMemcpyHostToDevice, s);
  if (!e && n) {
    k_ids_to_keys<<<blocks(n), BS, 0, s>>>(dIds, n, dc->keys.p, dc->dFlags);
    e = hipGetLastError();
  }
  if (!e) e = hipStreamSynchronize(s);
  (void)hipFree(dIds);
  if (e) return hip_fail(e, "dual contouring: cells", __FILE__, __LINE__);
  if (int r = check_flags(dc, s)) return r;
  // depth-0 ancestors appended (upsample.cpp:264)
  if (n && !fromPoints) {
    k_ancestor0<<<blocks(n), BS, 0, s>>>(dc->keys.p, n);
    MLOD_HIP(hipGetLastError());
  }
  size_t m = 0;
  if (int r = sort_unique(dc, dc->keys, (fromPoints ? 1 : 2) * (size_t)n, &m, s)) return r;
  // retreiveCorrectRoot (upsample.cpp:401-406) from the depth-0 cells
  uint32_t b[17];
  if (int r = depth_bounds(dc, b, s)) return r;
  std::vector<uint64_t> d0(b[1] - b[0]);
  if (!d0.empty()) MLOD_HIP(hipMemcpyAsync(d0.data(), dc->keys.p + b[0], d0.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {-INT_MAX, -INT_MAX, -INT_MAX};
  for (uint64_t k : d0) {
    const int p[3] = {kx(k), ky(k), kz(k)};
    for (int a = 0; a < 3; ++a) mn[a] = std::min(mn[a], p[a]), mx[a] = std::max(mx[a], p[a]);
  }
  int delta = 0;
  if (!d0.empty())
    for (int a = 0; a < 3; ++a) delta = std::max(delta, (int)std::ceil(float(std::log2(mx[a] - mn[a] + 1))));
  dc->diag = 1.f;
  dc->origin[0] = dc->origin[1] = dc->origin[2] = 0.f;
  if (fromPoints) delta = 0, cut = false;
  if (!d0.empty() && delta > 0) {
    k_shift<<<blocks(dc->keys.n), BS, 0, s>>>(dc->keys.p, (uint32_t)dc->keys.n, mn[0], mn[1], mn[2], delta, dc->dFlags);
    MLOD_HIP(hipGetLastError());
    if (int r = check_flags(dc, s)) return r;
    dc->diag = (float)(1 << delta);
    for (int a = 0; a < 3; ++a) dc->origin[a] = (float)mn[a];
  }
  // grade (upsample.cpp:67-93), then cells outside the root dropped
  if (int r = level_pass(dc, 8,
                         [](const uint64_t* k, uint32_t lo, uint32_t hi, uint64_t* out, hipStream_t st) {
                           k_grade_level<<<blocks(hi - lo), BS, 0, st>>>(k, lo, hi, out);
                         },
                         s))
    return r;
  if (int r = dc->flags8.reserve(dc->keys.n, 0, s)) return r;
  k_keep<<<blocks(dc->keys.n), BS, 0, s>>>(dc->keys.p, (uint32_t)dc->keys.n, kMaxDcDepth, dc->flags8.p);
  MLOD_HIP(hipGetLastError());
  if (int r = compact_keys(dc, s)) return r;
  // regularize_up (upsample.cpp:40-46)
  if (int r = level_pass(dc, 1,
                         [](const uint64_t* k, uint32_t lo, uint32_t hi, uint64_t* out, hipStream_t st) {
                           k_parent_level<<<blocks(hi - lo), BS, 0, st>>>(k, lo, hi, out);
                         },
                         s))
    return r;
  // cut (upsample.cpp:95-100)
  if (cut) {
    if (int r = dc->flags8.reserve(dc->keys.n, 0, s)) return r;
    k_keep<<<blocks(dc->keys.n), BS, 0, s>>>(dc->keys.p, (uint32_t)dc->keys.n, (int)maxDepth, dc->flags8.p);
    MLOD_HIP(hipGetLastError());
    if (int r = compact_keys(dc, s)) return r;
  }
  // regularize_siblings (upsample.cpp:48-65)
  {
    const size_t nk = dc->keys.n;
    if (int r = dc->keys.reserve(9 * nk, nk, s)) return r;
    if (nk) {
      k_siblings<<<blocks(nk), BS, 0, s>>>(dc->keys.p, (uint32_t)nk, dc->keys.p + nk);
      MLOD_HIP(hipGetLastError());
    }
    if (int r = sort_unique(dc, dc->keys, 9 * nk, &m, s)) return r;
  }
  // leaves and the finest depth
  const uint32_t nk = (uint32_t)dc->keys.n;
  if (int r = depth_bounds(dc, b, s)) return r;
  dc->D = 0;
  for (int d = 0; d <= kMaxDcDepth; ++d)
    if (b[d + 1] > b[d]) dc->D = d;
  if (int r = dc->leaf.reserve(nk, 0, s)) return r;
  if (nk) {
    k_leaf_flags<<<blocks(nk), BS, 0, s>>>(dc->keys.p, nk, dc->leaf.p);
    MLOD_HIP(hipGetLastError());
  }
  // leaf list (node indices, in node order)
  if (int r = dc->leaves.reserve(nk, 0, s)) return r;
  if (int r = dc->u32.reserve(nk, 0, s)) return r;
  if (nk) {
    size_t need = 0;
    MLOD_HIP(rocprim::select(nullptr, need, rocprim::counting_iterator<uint32_t>(0), (uint8_t*)nullptr, (uint32_t*)nullptr,
                             dc->dCount, nk, s));
    if (int r = ensure_tmp(dc, need)) return r;
    size_t bytes = dc->tmpBytes;
    MLOD_HIP(rocprim::select(dc->tmp, bytes, rocprim::counting_iterator<uint32_t>(0), dc->leaf.p, dc->leaves.p, dc->dCount,
                             nk, s));
    MLOD_HIP(hipMemcpyAsync(&dc->nLeaves, dc->dCount, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
  }
  dc->leaves.n = dc->nLeaves;
  // distinct leaf corners at depth D
  const size_t nc8 = 8 * (size_t)dc->nLeaves;
  if (int r = dc->ck.reserve(nc8, 0, s)) return r;
  if (dc->nLeaves) {
    k_leaf_corners<<<blocks(dc->nLeaves), BS, 0, s>>>(dc->keys.p, dc->leaves.p, dc->nLeaves, dc->D, dc->ck.p);
    MLOD_HIP(hipGetLastError());
  }
  if (int r = sort_unique(dc, dc->ck, nc8, &m, s)) return r;
  if (int r = dc->cf.reserve(dc->ck.n, 0, s)) return r;
  if (int r = dc->cpos.reserve(3 * dc->ck.n, 0, s)) return r;
  if (dc->ck.n) {
    float4 m00, m03;
    frame_matrix(dc, m00, m03);
    k_corner_pos<<<blocks(dc->ck.n), BS, 0, s>>>(dc->ck.p, (uint32_t)dc->ck.n, 1.0f / (float)(1 << dc->D), m00, m03,
                                                  dc->cpos.p);
    MLOD_HIP(hipGetLastError());
  }
  MLOD_HIP(hipStreamSynchronize(s));
  return MLOD_OK;
}

int extract(mlod_dc* dc, hipStream_t s) {
  dc->refined = false;
  const uint32_t nk = (uint32_t)dc->keys.n, nl = dc->nLeaves, nc = (uint32_t)dc->ck.n;
  if (int r = dc->tri.reserve((size_t)nl * 12 * 6, 0, s)) return r;
  if (int r = dc->used.reserve(nk, 0, s)) return r;
  if (int r = dc->vid.reserve((size_t)nk + 1, 0, s)) return r;
  if (nk) MLOD_HIP(hipMemsetAsync(dc->used.p, 0, nk, s));
  const DcView v{dc->keys.p, dc->leaf.p, nk, dc->ck.p, dc->cf.p, nc, dc->D};
  if (nl) {
    k_min_edges<<<blocks((uint64_t)nl * 12), BS, 0, s>>>(v, dc->leaves.p, nl, dc->tri.p, dc->used.p);
    MLOD_HIP(hipGetLastError());
  }
  // vertex ids: exclusive scan of used (node order)
  dc->nVerts = 0;
  if (nk) {
    size_t need = 0;
    MLOD_HIP(rocprim::exclusive_scan(nullptr, need, dc->used.p, dc->vid.p, 0u, (size_t)nk + 1, rocprim::plus<uint32_t>(), s));
    if (int r = ensure_tmp(dc, need)) return r;
    // scan nk + 1 elements so the last one is the total: used[nk] must read as 0
    if (int r = dc->flags8.reserve((size_t)nk + 1, 0, s)) return r;
    MLOD_HIP(hipMemcpyAsync(dc->flags8.p, dc->used.p, nk, hipMemcpyDeviceToDevice, s));
    MLOD_HIP(hipMemsetAsync(dc->flags8.p + nk, 0, 1, s));
    size_t bytes = dc->tmpBytes;
    MLOD_HIP(rocprim::exclusive_scan(dc->tmp, bytes, dc->flags8.p, dc->vid.p, 0u, (size_t)nk + 1, rocprim::plus<uint32_t>(), s));
    MLOD_HIP(hipMemcpyAsync(&dc->nVerts, dc->vid.p + nk, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
  }
  if (int r = dc->verts.reserve(3 * (size_t)dc->nVerts, 0, s)) return r;
  if (int r = dc->vframe.reserve(3 * (size_t)dc->nVerts, 0, s)) return r;
  if (int r = dc->vnode.reserve(dc->nVerts, 0, s)) return r;
  if (nk) {
    k_vertices<<<blocks(nk), BS, 0, s>>>(v, dc->used.p, dc->vid.p, dc->verts.p, dc->vnode.p);
    MLOD_HIP(hipGetLastError());
  }
  // triangles: compact the emitted slots (in (leaf, edge) order), then vertex ids
  dc->nTris = 0;
  const size_t slots = (size_t)nl * 24;
  if (slots) {
    if (3 * slots > 0xffffffffu) return fail(MLOD_ERR_UNSUPPORTED, "dual contouring: too many leaf edges");
    if (int r = dc->u32.reserve(3 * slots, 0, s)) return r;
    // a triangle is emitted when its first index is set; select whole triangles with a flag per index
    if (int r = dc->flags8.reserve(3 * slots, 0, s)) return r;
    struct Flag {
      const uint32_t* t;
      __device__ uint8_t operator()(uint32_t i) const { return t[i - i % 3] != 0xffffffffu ? 1 : 0; }
    };
    auto flagIt = rocprim::make_transform_iterator(rocprim::counting_iterator<uint32_t>(0), Flag{dc->tri.p});
    size_t need = 0;
    MLOD_HIP(rocprim::select(nullptr, need, dc->tri.p, flagIt, dc->u32.p, dc->dCount, 3 * slots, s));
    if (int r = ensure_tmp(dc, need)) return r;
    size_t bytes = dc->tmpBytes;
    MLOD_HIP(rocprim::select(dc->tmp, bytes, dc->tri.p, flagIt, dc->u32.p, dc->dCount, 3 * slots, s));
    uint32_t cnt = 0;
    MLOD_HIP(hipMemcpyAsync(&cnt, dc->dCount, sizeof(cnt), hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
    dc->nTris = cnt / 3;
    if (cnt) {
      k_tri_ids<<<blocks(cnt), BS, 0, s>>>(dc->u32.p, dc->nTris, dc->vid.p);
      MLOD_HIP(hipGetLastError());
    }
  }
  if (dc->nVerts) {
    float4 m00, m03;
    frame_matrix(dc, m00, m03);
    k_to_frame<<<blocks(dc->nVerts), BS, 0, s>>>(dc->verts.p, dc->nVerts, m00, m03, dc->vframe.p);
    MLOD_HIP(hipGetLastError());
  }
  MLOD_HIP(hipStreamSynchronize(s));
  dc->extracted = true;
  return MLOD_OK;
}

}  // namespace
}  // namespace mlod

using namespace mlod;

extern "C" {

int mlod_dc_create(mlod_dc** out) {
  if (!out) return fail(MLOD_ERR_INVALID, "out is NULL");
  mlod_dc* dc = new mlod_dc();
  if (hipMalloc((void**)&dc->dCount, 64) != hipSuccess || hipMalloc((void**)&dc->dCount64, 64) != hipSuccess ||
      hipMalloc((void**)&dc->dFlags, 64) != hipSuccess || hipMalloc((void**)&dc->dBounds, 32 * sizeof(uint32_t)) != hipSuccess) {
    mlod_dc_destroy(dc);
    return fail(MLOD_ERR_OOM, "dual contouring: counters");
  }
  *out = dc;
  return MLOD_OK;
}

int mlod_dc_destroy(mlod_dc* dc) {
  if (!dc) return MLOD_OK;
  (void)hipDeviceSynchronize();
  dc->keys.release(), dc->work.release(), dc->ck.release();
  dc->leaf.release(), dc->used.release(), dc->flags8.release();
  dc->leaves.release(), dc->vid.release(), dc->tri.release(), dc->vnode.release(), dc->u32.release();
  dc->cf.release(), dc->cpos.release(), dc->verts.release(), dc->vframe.release();
  dc->rv.release(), dc->rn.release(), dc->rt.release();
  void* p[] = {dc->tmp, dc->dCount, dc->dCount64, dc->dFlags, dc->dBounds};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete dc;
  return MLOD_OK;
}

int mlod_dc_build(mlod_dc* dc, const int32_t* ids4, uint32_t n, uint32_t max_depth, int32_t cut, void* stream) {
  if (!dc || !ids4) return fail(MLOD_ERR_INVALID, "NULL argument");
  if (n == 0) return fail(MLOD_ERR_INVALID, "dual contouring: no cells");
  return build(dc, ids4, n, max_depth, cut != 0, false, (hipStream_t)stream);
}

int mlod_dc_build_from_points(mlod_dc* dc, const int32_t* ids4, uint32_t n, void* stream) {
  if (!dc || !ids4) return fail(MLOD_ERR_INVALID, "NULL argument");
  if (n == 0) return fail(MLOD_ERR_INVALID, "dual contouring: no cells");
  return build(dc, ids4, n, 0, false, true, (hipStream_t)stream);
}

int mlod_dc_info(const mlod_dc* dc, double* out9) {
  if (!dc || !out9) return fail(MLOD_ERR_INVALID, "NULL argument");
  out9[0] = (double)dc->keys.n;
  out9[1] = dc->nLeaves;
  out9[2] = (double)dc->ck.n;
  out9[3] = dc->D;
  out9[4] = dc->diag;
  for (int a = 0; a < 3; ++a) out9[5 + a] = dc->origin[a];
  out9[8] = dc->extracted ? 1.0 : 0.0;
  return MLOD_OK;
}

int mlod_dc_nodes(const mlod_dc* dc, int32_t* ids4, uint8_t* is_leaf, void* stream) {
  if (!dc) return fail(MLOD_ERR_INVALID, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  const size_t n = dc->keys.n;
  if (ids4 && n) {
    std::vector<uint64_t> k(n);
    MLOD_HIP(hipMemcpyAsync(k.data(), dc->keys.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < n; ++i) ids4[4 * i] = kd(k[i]), ids4[4 * i + 1] = kx(k[i]), ids4[4 * i + 2] = ky(k[i]),
                                   ids4[4 * i + 3] = kz(k[i]);
  }
  if (is_leaf && n) {
    MLOD_HIP(hipMemcpyAsync(is_leaf, dc->leaf.p, n, hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
  }
  return MLOD_OK;
}

int mlod_dc_corners(const mlod_dc* dc, int32_t* xyz3, float* pos, void* stream) {
  if (!dc) return fail(MLOD_ERR_INVALID, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  const size_t n = dc->ck.n;
  if (xyz3 && n) {
    std::vector<uint64_t> k(n);
    MLOD_HIP(hipMemcpyAsync(k.data(), dc->ck.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < n; ++i)
      xyz3[3 * i] = (int32_t)(k[i] & 0x1fffffu), xyz3[3 * i + 1] = (int32_t)((k[i] >> 21) & 0x1fffffu),
                xyz3[3 * i + 2] = (int32_t)(k[i] >> 42);
  }
  if (pos && n) {
    MLOD_HIP(hipMemcpyAsync(pos, dc->cpos.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost, s));
    MLOD_HIP(hipStreamSynchronize(s));
  }
  return MLOD_OK;
}

int mlod_dc_project_corners(mlod_dc* dc, mlod_tree* t, void* stream) {
  if (!dc || !t) return fail(MLOD_ERR_INVALID, "NULL argument");
  if (!t->kernel_set) return fail(MLOD_ERR_STATE, "no kernel set (mlod_set_kernel)");
  hipStream_t s = (hipStream_t)stream;
  const uint32_t nc = (uint32_t)dc->ck.n;
  dc->extracted = false;
  if (!nc) {
    dc->haveValues = true;
    return MLOD_OK;
  }
  // APSS::project of the corners, one step, default step bound (scene.cpp:1371-1375); the
  // uploaded queries are set aside and restored
  const QueryState saved = t->qs;
  t->qs = QueryState{};
  int r = queries_set(t, dc->cpos.p, nc, true, s);
  if (!r)
    r = (t->mode == MLOD_MODE_KNN || t->mode == MLOD_MODE_BALL) ? launch_neighbors(t, nc, 1, s)
                                                                : launch_project(t, nc, 1, t->radius, s);
  if (!r) {
    k_corner_values<<<blocks(nc), BS, 0, s>>>(t->qs.q, t->qs.qn, t->qs.perm, nc, dc->cpos.p, dc->cf.p);
    const hipError_t e = hipGetLastError();
    if (e) r = hip_fail(e, "dual contouring: corner values", __FILE__, __LINE__);
  }
  (void)hipStreamSynchronize(s);
  queries_free(t);
  t->qs = saved;
  if (!r) dc->haveValues = true;
  return r;
}

int mlod_dc_set_corner_values(mlod_dc* dc, const float* f, void* stream) {
  if (!dc || (dc->ck.n && !f)) return fail(MLOD_ERR_INVALID, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  if (dc->ck.n) MLOD_HIP(hipMemcpyAsync(dc->cf.p, f, dc->ck.n * sizeof(float), hipMemcpyHostToDevice, s));
  MLOD_HIP(hipStreamSynchronize(s));
  dc->haveValues = true;
  dc->extracted = false;
  return MLOD_OK;
}

int mlod_dc_corner_values(const mlod_dc* dc, float* f, void* stream) {
  if (!dc || !f) return fail(MLOD_ERR_INVALID, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  if (dc->ck.n) MLOD_HIP(hipMemcpyAsync(f, dc->cf.p, dc->ck.n * sizeof(float), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  return MLOD_OK;
}

int mlod_dc_extract(mlod_dc* dc, uint32_t* n_verts, uint32_t* n_tris, void* stream) {
  if (!dc) return fail(MLOD_ERR_INVALID, "NULL argument");
  if (!dc->haveValues) return fail(MLOD_ERR_STATE, "no corner values (mlod_dc_project_corners or mlod_dc_set_corner_values)");
  if (int r = extract(dc, (hipStream_t)stream)) return r;
  if (n_verts) *n_verts = dc->nVerts;
  if (n_tris) *n_tris = dc->nTris;
  return MLOD_OK;
}

int mlod_dc_mesh(const mlod_dc* dc, float* verts, float* unit_verts, uint32_t* vert_node, uint32_t* tris, void* stream) {
  if (!dc) return fail(MLOD_ERR_INVALID, "NULL argument");
  if (!dc->extracted) return fail(MLOD_ERR_STATE, "no mesh (mlod_dc_extract)");
  hipStream_t s = (hipStream_t)stream;
  if (verts && dc->nVerts)
    MLOD_HIP(hipMemcpyAsync(verts, dc->vframe.p, 3 * (size_t)dc->nVerts * sizeof(float), hipMemcpyDeviceToHost, s));
  if (unit_verts && dc->nVerts)
    MLOD_HIP(hipMemcpyAsync(unit_verts, dc->verts.p, 3 * (size_t)dc->nVerts * sizeof(float), hipMemcpyDeviceToHost, s));
  if (vert_node && dc->nVerts)
    MLOD_HIP(hipMemcpyAsync(vert_node, dc->vnode.p, (size_t)dc->nVerts * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (tris && dc->nTris)
    MLOD_HIP(hipMemcpyAsync(tris, dc->u32.p, 3 * (size_t)dc->nTris * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  return MLOD_OK;
}

}  // extern "C"

namespace {
// temporaries of one mesh pass
struct MeshBufs {
  uint64_t* key = nullptr;
  uint64_t* skey = nullptr;
  uint32_t* val = nullptr;
  uint32_t* sval = nullptr;
  uint32_t* a = nullptr;
  uint32_t* b = nullptr;
  uint32_t* c = nullptr;
  uint8_t* f = nullptr;
  void* tmp = nullptr;
  void* tmp2 = nullptr;
  ~MeshBufs() {
    (void)hipDeviceSynchronize();
    void* p[] = {key, skey, val, sval, a, b, c, f, tmp, tmp2};
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
};

template <class T>
hipError_t dalloc(T** p, size_t n) {
  return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}

int check_tris(const uint32_t* tris, uint32_t nt, uint32_t nv) {
  for (size_t i = 0; i < 3 * (size_t)nt; ++i)
    if (tris[i] >= nv) return fail(MLOD_ERR_INVALID, "mesh: a triangle index is not a vertex");
  return MLOD_OK;
}

// sorted encounter keys (stable) of the triangles dTri for the mesh pass
int mesh_sort(MeshBufs& B, const uint32_t* dTri, uint32_t nt, int edges, hipStream_t s) {
  const size_t ne = 3 * (size_t)nt;
  MLOD_HIP(dalloc(&B.key, ne));
  MLOD_HIP(dalloc(&B.skey, ne));
  MLOD_HIP(dalloc(&B.val, ne));
  MLOD_HIP(dalloc(&B.sval, ne));
  if (!nt) return MLOD_OK;
  k_mesh_keys<<<blocks(ne), BS, 0, s>>>(dTri, nt, edges, B.key, B.val);
  MLOD_HIP(hipGetLastError());
  size_t need = 0;
  MLOD_HIP(rocprim::radix_sort_pairs(nullptr, need, B.key, B.skey, B.val, B.sval, ne, 0, 64, s));
  MLOD_HIP(hipMalloc(&B.tmp, need + 256));
  MLOD_HIP(rocprim::radix_sort_pairs(B.tmp, need, B.key, B.skey, B.val, B.sval, ne, 0, 64, s));
  return MLOD_OK;
}

// mesh::smoothMesh on device arrays: dP (nv x 3), dTri (nt x 3) -> dOut (nv x 3)
int mesh_smooth_dev(const float* dP, uint32_t nv, const uint32_t* dTri, uint32_t nt, float* dOut, hipStream_t s) {
  if (nv == 0) return MLOD_OK;
  MeshBufs B;
  if (int r = mesh_sort(B, dTri, nt, 0, s)) return r;
  k_smooth<<<blocks(nv), BS, 0, s>>>(dP, nv, dTri, B.skey, B.sval, (uint32_t)(3 * (size_t)nt), dOut);
  MLOD_HIP(hipGetLastError());
  MLOD_HIP(hipStreamSynchronize(s));
  return MLOD_OK;
}

// mesh::subdivideMesh on device arrays: dP (nv x 3), dTri (nt x 3) -> dOut (room for nv + 3 nt
// vertices; *outNv written), dOutTri (4 nt x 3)
int mesh_subdivide_dev(const float* dP, uint32_t nv, const uint32_t* dTri, uint32_t nt, float* dOut, uint32_t* dOutTri,
                       uint32_t* outNv, hipStream_t s) {
  *outNv = nv;
  if (nv) MLOD_HIP(hipMemcpyAsync(dOut, dP, 3 * (size_t)nv * sizeof(float), hipMemcpyDeviceToDevice, s));
  if (nt == 0) {
    MLOD_HIP(hipStreamSynchronize(s));
    return MLOD_OK;
  }
  const size_t ne = 3 * (size_t)nt;
  MeshBufs B;
  if (int r = mesh_sort(B, dTri, nt, 1, s)) return r;
  MLOD_HIP(dalloc(&B.a, ne + 1));
  MLOD_HIP(dalloc(&B.b, ne + 1));
  MLOD_HIP(dalloc(&B.c, ne));
  MLOD_HIP(dalloc(&B.f, ne + 1));
  // heads: a = head index per sorted position, f = first-encounter flags (encounter order)
  MLOD_HIP(hipMemsetAsync(B.f, 0, ne + 1, s));
  k_edge_heads<<<blocks(ne), BS, 0, s>>>(B.skey, B.sval, (uint32_t)ne, B.a, B.f);
  MLOD_HIP(hipGetLastError());
  // b = the head of each sorted position's run (max-scan of the head indices); c = rank of each
  // first encounter (exclusive scan of the flags)
  size_t n1 = 0, n2 = 0;
  MLOD_HIP(rocprim::inclusive_scan(nullptr, n1, B.a, B.b, ne, rocprim::maximum<uint32_t>(), s));
  MLOD_HIP(rocprim::exclusive_scan(nullptr, n2, B.f, B.c, 0u, ne, rocprim::plus<uint32_t>(), s));
  MLOD_HIP(hipMalloc(&B.tmp2, std::max(n1, n2) + 256));
  size_t bytes = std::max(n1, n2);
  MLOD_HIP(rocprim::inclusive_scan(B.tmp2, bytes, B.a, B.b, ne, rocprim::maximum<uint32_t>(), s));
  bytes = std::max(n1, n2);
  MLOD_HIP(rocprim::exclusive_scan(B.tmp2, bytes, B.f, B.c, 0u, ne, rocprim::plus<uint32_t>(), s));
  uint32_t lastRank = 0;
  uint8_t lastFlag = 0;
  MLOD_HIP(hipMemcpyAsync(&lastRank, B.c + ne - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipMemcpyAsync(&lastFlag, B.f + ne - 1, 1, hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  const uint32_t nEdges = lastRank + lastFlag;
  // B.val is free now: the new vertex of every encounter
  k_edge_verts<<<blocks(ne), BS, 0, s>>>(B.skey, B.sval, B.b, B.c, (uint32_t)ne, nv, dP, B.val, dOut);
  MLOD_HIP(hipGetLastError());
  k_subdiv_tris<<<blocks(nt), BS, 0, s>>>(dTri, nt, B.val, dOutTri);
  MLOD_HIP(hipGetLastError());
  MLOD_HIP(hipStreamSynchronize(s));
  *outNv = nv + nEdges;
  return MLOD_OK;
}

// the first n device points dV projected `steps` steps on t (its kernel, settings and mode,
// default step bound) in place, normals into dN; the tree's uploaded queries are set aside
int project_device(mlod_tree* t, float* dV, float* dN, uint32_t n, uint32_t steps, hipStream_t s) {
  if (!n) return MLOD_OK;
  const QueryState saved = t->qs;
  t->qs = QueryState{};
  int r = queries_set(t, dV, n, true, s);
  if (!r && steps)
    r = (t->mode == MLOD_MODE_KNN || t->mode == MLOD_MODE_BALL) ? launch_neighbors(t, n, steps, s)
                                                                : launch_project(t, n, steps, t->radius, s);
  if (!r) r = queries_scatter(t, n, dV, dN, s);
  (void)hipStreamSynchronize(s);
  queries_free(t);
  t->qs = saved;
  return r;
}
}  // namespace

extern "C" {

int mlod_mesh_smooth(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, float* out_verts, void* stream) {
  if (!verts || !out_verts || (nt && !tris)) return fail(MLOD_ERR_INVALID, "NULL argument");
  if (nt > 0x55555555u) return fail(MLOD_ERR_UNSUPPORTED, "mesh: too many triangles");
  if (int r = check_tris(tris, nt, nv)) return r;
  if (nv == 0) return MLOD_OK;
  hipStream_t s = (hipStream_t)stream;
  DVec<float> P, out;
  DVec<uint32_t> tri;
  int r = P.reserve(3 * (size_t)nv, 0, s);
  if (!r) r = out.reserve(3 * (size_t)nv, 0, s);
  if (!r) r = tri.reserve(3 * (size_t)nt, 0, s);
  hipError_t e = hipSuccess;
  if (!r) e = hipMemcpyAsync(P.p, verts, 3 * (size_t)nv * sizeof(float), hipMemcpyHostToDevice, s);
  if (!r && !e && nt) e = hipMemcpyAsync(tri.p, tris, 3 * (size_t)nt * sizeof(uint32_t), hipMemcpyHostToDevice, s);
  if (!r && e) r = hip_fail(e, "mesh smoothing", __FILE__, __LINE__);
  if (!r) r = mesh_smooth_dev(P.p, nv, tri.p, nt, out.p, s);
  if (!r) {
    e = hipMemcpyAsync(out_verts, out.p, 3 * (size_t)nv * sizeof(float), hipMemcpyDeviceToHost, s);
    if (!e) e = hipStreamSynchronize(s);
    if (e) r = hip_fail(e, "mesh smoothing", __FILE__, __LINE__);
  }
  (void)hipStreamSynchronize(s);
  P.release(), out.release(), tri.release();
  return r;
}

int mlod_mesh_subdivide(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, float* out_verts,
                        uint32_t out_cap, uint32_t* out_tris, uint32_t* out_nv, void* stream) {
  if (!verts || !out_verts || !out_nv || (nt && (!tris || !out_tris))) return fail(MLOD_ERR_INVALID, "NULL argument");
  if (nt > 0x15555555u) return fail(MLOD_ERR_UNSUPPORTED, "mesh: too many triangles");
  if ((uint64_t)nv + 3ull * nt > 0xffffffffull) return fail(MLOD_ERR_UNSUPPORTED, "mesh: too many vertices");
  if (out_cap < nv + 3 * nt) return fail(MLOD_ERR_INVALID, "mesh: out_verts holds fewer than nv + 3 nt vertices");
  if (int r = check_tris(tris, nt, nv)) return r;
  hipStream_t s = (hipStream_t)stream;
  std::memcpy(out_verts, verts, 3 * (size_t)nv * sizeof(float));
  *out_nv = nv;
  if (nt == 0) return MLOD_OK;
  DVec<float> P, out;
  DVec<uint32_t> tri, otri;
  int r = P.reserve(3 * (size_t)nv, 0, s);
  if (!r) r = out.reserve(3 * ((size_t)nv + 3 * (size_t)nt), 0, s);
  if (!r) r = tri.reserve(3 * (size_t)nt, 0, s);
  if (!r) r = otri.reserve(12 * (size_t)nt, 0, s);
  hipError_t e = hipSuccess;
  if (!r) e = hipMemcpyAsync(P.p, verts, 3 * (size_t)nv * sizeof(float), hipMemcpyHostToDevice, s);
  if (!r && !e) e = hipMemcpyAsync(tri.p, tris, 3 * (size_t)nt * sizeof(uint32_t), hipMemcpyHostToDevice, s);
  if (!r && e) r = hip_fail(e, "mesh subdivision", __FILE__, __LINE__);
  uint32_t nOut = nv;
  if (!r) r = mesh_subdivide_dev(P.p, nv, tri.p, nt, out.p, otri.p, &nOut, s);
  if (!r) {
    e = hipMemcpyAsync(out_verts, out.p, 3 * (size_t)nOut * sizeof(float), hipMemcpyDeviceToHost, s);
    if (!e) e = hipMemcpyAsync(out_tris, otri.p, 12 * (size_t)nt * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (!e) e = hipStreamSynchronize(s);
    if (e) r = hip_fail(e, "mesh subdivision", __FILE__, __LINE__);
  }
  (void)hipStreamSynchronize(s);
  P.release(), out.release(), tri.release(), otri.release();
  if (!r) *out_nv = nOut;
  return r;
}

int mlod_dc_refine(mlod_dc* dc, mlod_tree* t, uint32_t steps, uint32_t rounds, uint32_t* n_verts, uint32_t* n_tris,
                   void* stream) {
  if (!dc || !t) return fail(MLOD_ERR_INVALID, "NULL argument");
  if (!dc->extracted) return fail(MLOD_ERR_STATE, "no mesh (mlod_dc_extract)");
  if (steps && !t->kernel_set) return fail(MLOD_ERR_STATE, "no kernel set (mlod_set_kernel)");
  hipStream_t s = (hipStream_t)stream;
  dc->refined = false;
  // sizes after the rounds: vertices grow by at most 3 per triangle, triangles 4x per round
  uint64_t nt = dc->nTris, nvMax = dc->nVerts;
  for (uint32_t r = 0; r < rounds && nt; ++r) nvMax += 3 * nt, nt *= 4;
  if (nt > 0x15555555ull || nvMax > 0xffffffffull) return fail(MLOD_ERR_UNSUPPORTED, "mesh: too many triangles");
  DVec<float>& V = dc->rv;
  DVec<float>& N = dc->rn;
  DVec<uint32_t>& F = dc->rt;
  DVec<float> V2;
  DVec<uint32_t> F2;
  int r = V.reserve(3 * (size_t)nvMax, 0, s);
  if (!r) r = N.reserve(3 * (size_t)nvMax, 0, s);
  if (!r) r = F.reserve(3 * (size_t)nt, 0, s);
  if (!r) r = V2.reserve(3 * (size_t)nvMax, 0, s);
  if (!r) r = F2.reserve(3 * (size_t)nt, 0, s);
  uint32_t nv = dc->nVerts, ntri = dc->nTris;
  if (!r && nv) {
    hipError_t e = hipMemcpyAsync(V.p, dc->vframe.p, 3 * (size_t)nv * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (!e && ntri) e = hipMemcpyAsync(F.p, dc->u32.p, 3 * (size_t)ntri * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (!e) e = hipMemsetAsync(N.p, 0, 3 * (size_t)nv * sizeof(float), s);
    if (e) r = hip_fail(e, "mesh refinement", __FILE__, __LINE__);
  }
  // scene.cpp:1417-1452: the vertices projected, then per round smoothing, subdivision and
  // projection again
  if (!r && steps) r = project_device(t, V.p, N.p, nv, steps, s);
  for (uint32_t k = 0; !r && k < rounds && ntri; ++k) {
    r = mesh_smooth_dev(V.p, nv, F.p, ntri, V2.p, s);
    uint32_t nv2 = nv;
    if (!r) r = mesh_subdivide_dev(V2.p, nv, F.p, ntri, V.p, F2.p, &nv2, s);
    if (!r) {
      std::swap(F, F2);
      nv = nv2, ntri *= 4;
      const hipError_t e = hipMemsetAsync(N.p, 0, 3 * (size_t)nv * sizeof(float), s);
      if (e) r = hip_fail(e, "mesh refinement", __FILE__, __LINE__);
    }
    if (!r && steps) r = project_device(t, V.p, N.p, nv, steps, s);
  }
  (void)hipStreamSynchronize(s);
  V2.release(), F2.release();
  if (r) return r;
  dc->nRV = nv, dc->nRT = ntri, dc->refined = true;
  if (n_verts) *n_verts = nv;
  if (n_tris) *n_tris = ntri;
  return MLOD_OK;
}

int mlod_dc_refined_mesh(const mlod_dc* dc, float* verts, float* normals, uint32_t* tris, void* stream) {
  if (!dc) return fail(MLOD_ERR_INVALID, "NULL argument");
  if (!dc->refined) return fail(MLOD_ERR_STATE, "no refined mesh (mlod_dc_refine)");
  hipStream_t s = (hipStream_t)stream;
  if (verts && dc->nRV) MLOD_HIP(hipMemcpyAsync(verts, dc->rv.p, 3 * (size_t)dc->nRV * sizeof(float), hipMemcpyDeviceToHost, s));
  if (normals && dc->nRV)
    MLOD_HIP(hipMemcpyAsync(normals, dc->rn.p, 3 * (size_t)dc->nRV * sizeof(float), hipMemcpyDeviceToHost, s));
  if (tris && dc->nRT) MLOD_HIP(hipMemcpyAsync(tris, dc->rt.p, 3 * (size_t)dc->nRT * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  MLOD_HIP(hipStreamSynchronize(s));
  return MLOD_OK;
}

}  // extern "C"

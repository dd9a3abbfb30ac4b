// neighbors.hip — kNN and ball APSS on the GPU (gfx950).
//
// Restates APSSOctree::projectAPSSCPU (sources/apss/CPUUtils.cpp:690-737), which the
// reference runs on the CPU only, over its kd-tree (kdTreeNN.h:416-493):
//   per step: the k nearest input points of the current position (kdTree::knearest) or all
//   points with d^2 <= r^2 (getInBall; r < 0 => 3 sigmas[0]); unit weights w(p_i, q) on the
//   raw points and normals; the sphere fit of projectOnAPSSStats called with q aliased to
//   the output (CPUUtils.cpp:735), so no step clamp applies and a NaN fit leaves a NaN
//   position (normal zeroed); an empty ball returns the step's starting query with a zero
//   normal and ends the iterations (CPUUtils.cpp:711-716).
// Neighbours are found by walking this library's octree with box-distance pruning instead of
// a kd-tree; distances use the kd-tree's arithmetic (float differences, squares summed in
// float, compared as double), so the neighbour sets are the same; the k nearest are ordered
// by (distance, point index).
//
// kNN: one lane per query.  Per depth level the lane keeps in LDS the first child and the
// children still to visit in increasing box distance (so the search reaches the query's own
// leaves first and the k-th distance bound tightens early); small subtrees are scanned as
// contiguous point ranges.  Candidates are kept in registers as a sorted array of 64-bit keys
// (distance bits << 32 | sorted point index).  Ball: one walk per wavefront over the union of
// its lanes' balls (k_ball_wave).
#include <cfloat>
#include <cmath>

#include "mlod_internal.h"

namespace mlod {
namespace {

constexpr int NBS = 128;
constexpr int kLevels = kMaxDepth;

#define LAUNCH_CHECK()                                                           \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) return hip_fail(e_, "kernel launch", __FILE__, __LINE__); \
  } while (0)

template <class T>
struct V3 {
  T x, y, z;
};

// kdTreeNN distance arithmetic (kdTreeNN.h:421, 470): ((dx*dx + dy*dy) + dz*dz) in float.
// This file is compiled without multiply-add contraction.
__device__ __forceinline__ float kd_d2(float4 p, V3<float> q) {
  const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
  return (dx * dx + dy * dy) + dz * dz;
}

// Squared distance from q to a node's closed box (a lower bound for all its points).
__device__ __forceinline__ float box_d2(const float* __restrict__ box, uint32_t c, V3<float> q) {
  const float2 a = *reinterpret_cast<const float2*>(box + 6 * (size_t)c);
  const float2 b = *reinterpret_cast<const float2*>(box + 6 * (size_t)c + 2);
  const float2 e = *reinterpret_cast<const float2*>(box + 6 * (size_t)c + 4);
  // a = (lo.x, lo.y), b = (lo.z, hi.x), e = (hi.y, hi.z)
  const float dx = fmaxf(fmaxf(a.x - q.x, q.x - b.y), 0.f);
  const float dy = fmaxf(fmaxf(a.y - q.y, q.y - e.x), 0.f);
  const float dz = fmaxf(fmaxf(b.x - q.z, q.z - e.y), 0.f);
  return (dx * dx + dy * dy) + dz * dz;
}

// Bound slack: box distances are float-rounded, so a node is kept while its bound is within
// a relative 1e-5 of the search radius (a superset of the exact test).
constexpr float kSlack = 1.0001f;

// kNN: subtrees of at most this many points are scanned rather than walked.
constexpr uint32_t kScanPointsKnn = 8;

// Weighted sums of the neighbours (CPUUtils.cpp:719-731): unit area, raw p and n.
struct SumF {  // plain FP32
  float w, px, py, pz, nx, ny, nz, pn, pp;
};
struct KSum {
  float s, c;
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double val() const { return (double)s - (double)c; }
};
struct SumC {  // compensated, query frame
  KSum w, px, py, pz, nx, ny, nz, pn, pp;
};
struct SumD {
  double w, px, py, pz, nx, ny, nz, pn, pp;
};

template <int KT>
__device__ __forceinline__ float weight32(const KParams& k, float d2) {
  if constexpr (KT == MLOD_KERNEL_GAUSSIAN_MULTIPLE) {
    float w = k.b[0] * __builtin_amdgcn_exp2f(d2 * k.a[0]);
    for (int s = 1; s < k.ns; ++s) w = fmaf(k.b[s], __builtin_amdgcn_exp2f(d2 * k.a[s]), w);
    return w == 0.f ? k.c : w;
  } else if constexpr (KT == MLOD_KERNEL_GAUSSIAN) {
    return __builtin_amdgcn_exp2f(d2 * k.a[0]) + k.c;
  } else {
    const float base = k.cp / (d2 + k.c);
    if (k.eint > 0) {
      float r = 1.f, b = base;
      for (int e = k.eint; e; e >>= 1) {
        if (e & 1) r *= b;
        b *= b;
      }
      return r;
    }
    return __builtin_amdgcn_exp2f(k.e * __builtin_amdgcn_logf(base));
  }
}

template <int KT>
__device__ __forceinline__ double weight64(const KParams& k, double d2) {
  if constexpr (KT == MLOD_KERNEL_GAUSSIAN_MULTIPLE) {
    double w = 0.0;
    for (int s = 0; s < k.ns; ++s) w += exp(d2 * k.ad[s]) * k.bd[s];
    return w == 0.0 ? (double)k.c : w;
  } else if constexpr (KT == MLOD_KERNEL_GAUSSIAN) {
    return exp(d2 * k.ad[0]) + (double)k.c;
  } else {
    return pow((double)k.cp, (double)k.e) / pow(d2 + (double)k.c, (double)k.e);
  }
}

template <int FORM>
struct Sums;
template <> struct Sums<0> { using S = SumF; };
template <> struct Sums<1> { using S = SumC; };
template <> struct Sums<2> { using S = SumD; };

template <int FORM, int KT>
__device__ __forceinline__ void add_nb(typename Sums<FORM>::S& s, const KParams& kp, float4 P, float4 N, V3<double> qd,
                                       V3<float> qf) {
  if constexpr (FORM == 0) {
    const float dx = qf.x - P.x, dy = qf.y - P.y, dz = qf.z - P.z;
    const float w = weight32<KT>(kp, (dx * dx + dy * dy) + dz * dz);
    s.pn += w * ((P.x * N.x + P.y * N.y) + P.z * N.z);
    s.pp += w * ((P.x * P.x + P.y * P.y) + P.z * P.z);
    s.px += w * P.x;
    s.py += w * P.y;
    s.pz += w * P.z;
    s.nx += w * N.x;
    s.ny += w * N.y;
    s.nz += w * N.z;
    s.w += w;
  } else if constexpr (FORM == 1) {
    const float dx = P.x - qf.x, dy = P.y - qf.y, dz = P.z - qf.z;
    const float d2 = (dx * dx + dy * dy) + dz * dz;
    const float w = weight32<KT>(kp, d2);
    s.w.add(w);
    s.px.add(w * dx);
    s.py.add(w * dy);
    s.pz.add(w * dz);
    s.nx.add(w * N.x);
    s.ny.add(w * N.y);
    s.nz.add(w * N.z);
    s.pn.add(w * ((dx * N.x + dy * N.y) + dz * N.z));
    s.pp.add(w * d2);
  } else {
    const double x = P.x, y = P.y, z = P.z;
    const double dx = qd.x - x, dy = qd.y - y, dz = qd.z - z;
    const double w = weight64<KT>(kp, dx * dx + dy * dy + dz * dz);
    s.pn += w * ((x * N.x + y * N.y) + z * N.z);
    s.pp += w * ((x * x + y * y) + z * z);
    s.px += w * x;
    s.py += w * y;
    s.pz += w * z;
    s.nx += w * N.x;
    s.ny += w * N.y;
    s.nz += w * N.z;
    s.w += w;
  }
}

template <class T>
__device__ __forceinline__ T dot3(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// projectOnAPSSStats with q aliased to the output (CPUUtils.cpp:735, 739-782): plane
// threshold 1e-12, normalize(0) = NaN, NaN positions kept with a zero normal, no step clamp.
// kind: 0 sphere, 1 plane, 2 invalid (NaN position kept, zero normal)
template <class T>
__device__ __forceinline__ void fit_aliased(T sw, V3<T> sp, V3<T> sn, T spn, T spp, V3<T> q, V3<T>& outP, V3<T>& outN,
                                            float& kind) {
  const T A = spn - dot3(sp, V3<T>{sn.x / sw, sn.y / sw, sn.z / sw});
  const T B = spp - dot3(sp, V3<T>{sp.x / sw, sp.y / sw, sp.z / sw});
  const T u4 = (T)((0.5 * (double)A) / (double)B);
  const T t2 = T(2) * u4;
  const V3<T> u{(sn.x - t2 * sp.x) / sw, (sn.y - t2 * sp.y) / sw, (sn.z - t2 * sp.z) / sw};
  const T u0 = -(dot3(sp, u) + u4 * spp) / sw;
  V3<T> p = q, n;
  kind = fabs((double)u4) < 0.000000000001 ? 1.f : 0.f;
  if (fabs((double)u4) < 0.000000000001) {
    const V3<T> nn{-u.x, -u.y, -u.z};
    const T lambda = (u0 - dot3(p, nn)) / dot3(nn, nn);
    p = V3<T>{p.x + lambda * nn.x, p.y + lambda * nn.y, p.z + lambda * nn.z};
    const T inv = T(1) / sqrt(dot3(sn, sn));
    n = V3<T>{sn.x * inv, sn.y * inv, sn.z * inv};
  } else {
    const T m2 = T(-2) * u4;
    const V3<T> c{u.x / m2, u.y / m2, u.z / m2};
    const T v = dot3(c, c) - u0 / u4;
    const T r = sqrt(v > T(0) ? v : T(0));
    const V3<T> d{p.x - c.x, p.y - c.y, p.z - c.z};
    const T inv = T(1) / sqrt(dot3(d, d));
    p = V3<T>{c.x + r * d.x * inv, c.y + r * d.y * inv, c.z + r * d.z * inv};
    const V3<T> g{u.x + t2 * p.x, u.y + t2 * p.y, u.z + t2 * p.z};
    const T ig = T(1) / sqrt(dot3(g, g));
    n = V3<T>{g.x * ig, g.y * ig, g.z * ig};
  }
  if (isnan(p.x) || isnan(p.y) || isnan(p.z)) {
    n = V3<T>{0, 0, 0};
    kind = 2.f;
  }
  outP = p;
  outN = n;
}

template <int FORM>
__device__ __forceinline__ void fit_step(const typename Sums<FORM>::S& s, V3<double>& qd, V3<float> qf, float4& nout) {
  if constexpr (FORM == 0) {
    V3<float> p, n;
    float kind;
    fit_aliased<float>(s.w, V3<float>{s.px, s.py, s.pz}, V3<float>{s.nx, s.ny, s.nz}, s.pn, s.pp, qf, p, n, kind);
    qd = V3<double>{p.x, p.y, p.z};
    nout = make_float4(n.x, n.y, n.z, kind);
  } else if constexpr (FORM == 1) {
    V3<double> p, n;
    float kind;
    fit_aliased<double>(s.w.val(), V3<double>{s.px.val(), s.py.val(), s.pz.val()},
                        V3<double>{s.nx.val(), s.ny.val(), s.nz.val()}, s.pn.val(), s.pp.val(), V3<double>{0, 0, 0}, p, n,
                        kind);
    qd = V3<double>{(double)qf.x + p.x, (double)qf.y + p.y, (double)qf.z + p.z};
    nout = make_float4((float)n.x, (float)n.y, (float)n.z, kind);
  } else {
    V3<double> p, n;
    float kind;
    fit_aliased<double>(s.w, V3<double>{s.px, s.py, s.pz}, V3<double>{s.nx, s.ny, s.nz}, s.pn, s.pp, qd, p, n, kind);
    qd = p;
    nout = make_float4((float)n.x, (float)n.y, (float)n.z, kind);
  }
}

// Sorted (ascending) candidate keys of the k nearest, in registers.  Real keys are
// (distance bits << 32) | (sorted point index + 1); the first KMAX - k slots hold the
// sentinel 0, smaller than every real key and never displaced, so the k-th smallest real key
// is always the last slot (a static register index).
template <int KMAX>
struct TopK {
  uint64_t key[KMAX];
  __device__ __forceinline__ void clear(uint32_t k) {
#pragma unroll
    for (int j = 0; j < KMAX; ++j) key[j] = (uint32_t)j < KMAX - k ? 0ull : ~0ull;
  }
  __device__ __forceinline__ void insert(uint64_t x) {
    if (x >= key[KMAX - 1]) return;
#pragma unroll
    for (int j = KMAX - 1; j > 0; --j) {
      const uint64_t lo = key[j - 1] < x ? x : key[j - 1];  // max(key[j-1], x)
      key[j] = key[j] < lo ? key[j] : lo;                   // min(key[j], max(key[j-1], x))
    }
    key[0] = key[0] < x ? key[0] : x;
  }
  // removes and returns the smallest key (register shift, static indices only)
  __device__ __forceinline__ uint64_t pop_front() {
    const uint64_t v = key[0];
#pragma unroll
    for (int j = 0; j < KMAX - 1; ++j) key[j] = key[j + 1];
    key[KMAX - 1] = ~0ull;
    return v;
  }
  __device__ __forceinline__ uint64_t kth() const { return key[KMAX - 1]; }
};

template <int FORM, int KT, int KMAX>
// The 32 sorted 64-bit keys take 64 VGPRs; the FP32 forms are held to three waves per SIMD (the
// compensated form then spills ~64 B per lane, cheaper than the lost wave), the FP64 form keeps two.
__global__ __launch_bounds__(NBS) __attribute__((amdgpu_waves_per_eu(FORM == 2 ? 1 : 3)))
void k_knn(TreeView tv, const float* __restrict__ box, KParams kp,
                                             float4* __restrict__ q, float4* __restrict__ qn,
                                             const uint32_t* __restrict__ perm, uint32_t m, uint32_t limit, uint32_t nIter,
                                             uint32_t k) {
  __shared__ uint32_t sFirst[kLevels][NBS];
  __shared__ uint32_t sOrder[kLevels][NBS];  // children still to visit, nearest first, 4 bits each
  __shared__ uint8_t sLeft[kLevels][NBS];    // how many
  const int tid = threadIdx.x;
  const uint32_t i = xcd_block() * NBS + tid;
  if (i >= m) return;
  if (limit < m && perm[i] >= limit) return;
  const float4 q4 = q[i];
  V3<double> qd{q4.x, q4.y, q4.z};
  float4 nout = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t it = 0; it < nIter; ++it) {
    const V3<float> qf{(float)qd.x, (float)qd.y, (float)qd.z};
    TopK<KMAX> top;
    top.clear(k);
    float bound = INFINITY;  // squared distance of the current k-th candidate
    const uint4 h0 = tv.hdr[0];
    int lvl = -1;
    uint32_t cur = 0;  // node to expand next (root first)
    bool have = true;
    while (true) {
      if (have) {
        const uint4 h = cur == 0 ? h0 : tv.hdr[cur];
        const uint2 sp = tv.span[cur];
        if (hdr_leaf(h.y) || sp.y - sp.x <= kScanPointsKnn) {
          // a leaf or a small subtree: every point is a candidate (the k nearest do not depend
          // on the order they are offered in)
          for (uint32_t j = sp.x; j < sp.y; ++j) {
            const float d2 = kd_d2(tv.pts[j], qf);
            top.insert(((uint64_t)__float_as_uint(d2) << 32) | (j + 1u));
          }
          const uint64_t kk = top.kth();
          bound = kk == ~0ull ? INFINITY : __uint_as_float((uint32_t)(kk >> 32));
        } else {
          // children whose box reaches the bound, visited nearest box first, so that the
          // first leaves reached are the query's own and the bound tightens at once
          const uint32_t first = h.x, nc = hdr_count(h.y);
          // keys: the box distance's bits (non-negative floats order as integers) with the child
          // slot in the low 3 bits, children beyond the bound last; an 8-input sorting network
          // (19 compare-exchanges) orders them.  The visiting order only steers the search: the
          // k nearest do not depend on it
          uint32_t key[8], cnt = 0;
#pragma unroll
          for (uint32_t c = 0; c < 8; ++c) {
            const float d = c < nc ? box_d2(box, first + c, qf) : INFINITY;
            const bool in = c < nc && d <= bound * kSlack;
            key[c] = in ? (__float_as_uint(d) & ~7u) | c : 0xffffffffu;
            cnt += in ? 1u : 0u;
          }
          const auto cx = [&](int a, int b) {
            const uint32_t lo = min(key[a], key[b]), hi = max(key[a], key[b]);
            key[a] = lo, key[b] = hi;
          };
          cx(0, 2), cx(1, 3), cx(4, 6), cx(5, 7), cx(0, 4), cx(1, 5), cx(2, 6), cx(3, 7), cx(0, 1), cx(2, 3);
          cx(4, 5), cx(6, 7), cx(2, 4), cx(3, 5), cx(1, 4), cx(3, 6), cx(1, 2), cx(3, 4), cx(5, 6);
          uint32_t order = 0;
#pragma unroll
          for (uint32_t r = 0; r < 8; ++r) order |= (key[r] & 7u) << (4u * r);
          ++lvl;
          sFirst[lvl][tid] = first;
          sOrder[lvl][tid] = order;
          sLeft[lvl][tid] = (uint8_t)cnt;
        }
        have = false;
      }
      if (lvl < 0) break;
      const uint32_t left = sLeft[lvl][tid];
      if (left == 0) {
        --lvl;
        continue;
      }
      const uint32_t order = sOrder[lvl][tid];
      sOrder[lvl][tid] = order >> 4;
      sLeft[lvl][tid] = (uint8_t)(left - 1u);
      const uint32_t c = sFirst[lvl][tid] + (order & 15u);
      if (box_d2(box, c, qf) <= bound * kSlack) {  // the bound may have shrunk since the push
        cur = c;
        have = true;
      }
    }
    typename Sums<FORM>::S s{};
#pragma unroll 1
    for (uint32_t j = 0; j < KMAX; ++j) {
      const uint64_t key = top.pop_front();
      if (key == 0ull) continue;  // sentinel
      if (key == ~0ull) break;    // fewer than k points in the cloud
      const uint32_t pj = (uint32_t)key - 1u;
      add_nb<FORM, KT>(s, kp, tv.pts[pj], tv.nrm[pj], qd, qf);
    }
    fit_step<FORM>(s, qd, qf, nout);
  }
  q[i] = make_float4((float)qd.x, (float)qd.y, (float)qd.z, 0.f);
  qn[i] = nout;
}

// Ball APSS, one wavefront at a time.  The 64 lanes' queries are Morton neighbours, so their
// balls overlap almost entirely: the wavefront walks the octree once over the union of its
// balls (node boxes against the lanes' query box grown by r), takes the points of every
// reached subtree that is small enough as a contiguous range, loads each range 64 points at a
// time with one coalesced load per lane, and broadcasts the points lane to lane; every lane
// sums the points inside its own ball.  Ranges come in increasing point order, so each lane
// sums exactly the points inside its ball, in increasing point order (the order of a walk over
// its own ball alone), whatever the other lanes' queries.
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float bcast(float v, uint32_t lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)lane));
}

// squared distance between a node's box and the box [lo, hi]
__device__ __forceinline__ float box_box_d2(const float* __restrict__ box, uint32_t c, V3<float> lo, V3<float> hi) {
  const float2 a = *reinterpret_cast<const float2*>(box + 6 * (size_t)c);
  const float2 b = *reinterpret_cast<const float2*>(box + 6 * (size_t)c + 2);
  const float2 e = *reinterpret_cast<const float2*>(box + 6 * (size_t)c + 4);
  const float dx = fmaxf(fmaxf(a.x - hi.x, lo.x - b.y), 0.f);
  const float dy = fmaxf(fmaxf(a.y - hi.y, lo.y - e.x), 0.f);
  const float dz = fmaxf(fmaxf(b.x - hi.z, lo.z - e.y), 0.f);
  return (dx * dx + dy * dy) + dz * dz;
}

constexpr uint32_t kWaveScanPoints = 256;  // subtrees up to this size are taken as one range

template <int FORM, int KT>
__global__ __launch_bounds__(NBS) void k_ball_wave(TreeView tv, const float* __restrict__ box, KParams kp,
                                                   float4* __restrict__ q, float4* __restrict__ qn,
                                                   const uint32_t* __restrict__ perm, uint32_t m, uint32_t limit,
                                                   uint32_t nIter, float r2f) {
  constexpr int kWaves = NBS / 64;
  __shared__ uint32_t sFirst[kLevels][kWaves];
  __shared__ uint32_t sMask[kLevels][kWaves];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint32_t i = xcd_block() * NBS + tid;
  const uint32_t i0 = i - (uint32_t)lane;  // the wavefront's first slot
  if (i0 >= m) return;                      // whole wavefront past the end
  const bool mine = i < m && !(limit < m && perm[i] >= limit);
  // lanes without a query of their own follow the wavefront's first query (never stored)
  const float4 q4 = q[i < m ? i : i0];
  V3<double> qd{q4.x, q4.y, q4.z};
  float4 nout = make_float4(0.f, 0.f, 0.f, 0.f);
  bool done = !mine;
  const float r = sqrtf(r2f) * kSlack;
  for (uint32_t it = 0; it < nIter; ++it) {
    const V3<float> qf{(float)qd.x, (float)qd.y, (float)qd.z};
    // the union of the live lanes' balls, bounded by their query box grown by r
    const float big = 3.0e38f;
    const V3<float> qlo{wave_min(done ? big : qf.x) - r, wave_min(done ? big : qf.y) - r, wave_min(done ? big : qf.z) - r};
    const V3<float> qhi{wave_max(done ? -big : qf.x) + r, wave_max(done ? -big : qf.y) + r, wave_max(done ? -big : qf.z) + r};
    if (qlo.x > qhi.x) break;  // every lane finished (uniform)
    typename Sums<FORM>::S s{};
    uint32_t cnt = 0;
    int lvl = -1;
    uint32_t cur = 0;
    bool have = true;
    while (true) {
      if (have) {
        const uint4 h = tv.hdr[cur];
        const uint2 sp = tv.span[cur];
        if (hdr_leaf(h.y) || sp.y - sp.x <= kWaveScanPoints) {
          // the range's points, 64 at a time: one coalesced load per lane, then each point
          // is broadcast and tested by every lane
          for (uint32_t base = sp.x; base < sp.y; base += 64) {
            const uint32_t nb = min(64u, sp.y - base);
            const uint32_t j = base + (uint32_t)lane;
            const float4 P = (uint32_t)lane < nb ? tv.pts[j] : make_float4(0.f, 0.f, 0.f, 0.f);
            const float4 N = (uint32_t)lane < nb ? tv.nrm[j] : make_float4(0.f, 0.f, 0.f, 0.f);
            for (uint32_t t = 0; t < nb; ++t) {
              const float4 Pt = make_float4(bcast(P.x, t), bcast(P.y, t), bcast(P.z, t), bcast(P.w, t));
              if (!done && kd_d2(Pt, qf) <= r2f) {
                const float4 Nt = make_float4(bcast(N.x, t), bcast(N.y, t), bcast(N.z, t), 0.f);
                add_nb<FORM, KT>(s, kp, Pt, Nt, qd, qf);
                ++cnt;
              }
            }
          }
        } else {
          const uint32_t first = __builtin_amdgcn_readfirstlane(h.x), nc = hdr_count(h.y);
          uint32_t mask = 0;
          for (uint32_t c = 0; c < nc; ++c)
            if (box_box_d2(box, first + c, qlo, qhi) <= 0.f) mask |= 1u << c;
          ++lvl;
          if (lane == 0) {
            sFirst[lvl][wv] = first;
            sMask[lvl][wv] = mask;
          }
        }
        have = false;
      }
      if (lvl < 0) break;
      const uint32_t mask = __builtin_amdgcn_readfirstlane(sMask[lvl][wv]);
      if (mask == 0) {
        --lvl;
        continue;
      }
      if (lane == 0) sMask[lvl][wv] = mask & (mask - 1u);
      cur = __builtin_amdgcn_readfirstlane(sFirst[lvl][wv]) + (uint32_t)__builtin_ctz(mask);
      have = true;
    }
    if (!done) {
      if (cnt == 0) {  // empty ball: the starting query, zero normal (CPUUtils.cpp:711-716)
        qd = V3<double>{q4.x, q4.y, q4.z};
        nout = make_float4(0.f, 0.f, 0.f, 0.f);
        done = true;
      } else {
        fit_step<FORM>(s, qd, qf, nout);
      }
    }
  }
  if (mine) {
    q[i] = make_float4((float)qd.x, (float)qd.y, (float)qd.z, 0.f);
    qn[i] = nout;
  }
}

template <int FORM, int KT>
int launch_form(mlod_tree* t, const KParams& kp, uint32_t limit, uint32_t nIter, hipStream_t s) {
  const uint32_t m = t->qs.m;
  const uint32_t grid = (m + NBS - 1) / NBS;
  if (t->mode == MLOD_MODE_KNN) {
    const uint32_t k = t->knn < t->n ? t->knn : t->n;
    if (k <= 8)
      k_knn<FORM, KT, 8><<<grid, NBS, 0, s>>>(t->view(), t->box, kp, t->qs.q, t->qs.qn, t->qs.perm, m, limit, nIter, k);
    else if (k <= 16)
      k_knn<FORM, KT, 16><<<grid, NBS, 0, s>>>(t->view(), t->box, kp, t->qs.q, t->qs.qn, t->qs.perm, m, limit, nIter, k);
    else if (k <= 32)
      k_knn<FORM, KT, 32><<<grid, NBS, 0, s>>>(t->view(), t->box, kp, t->qs.q, t->qs.qn, t->qs.perm, m, limit, nIter, k);
    else if (k <= 64)
      k_knn<FORM, KT, 64><<<grid, NBS, 0, s>>>(t->view(), t->box, kp, t->qs.q, t->qs.qn, t->qs.perm, m, limit, nIter, k);
    else
      return fail(MLOD_ERR_UNSUPPORTED, "kNN supports k <= 64");
  } else {
    // getInBall(q, r) with r = sigmas[0] * 3 when r < 0 (CPUUtils.cpp:706-709), d^2 <= r*r in
    // double; r2f is the largest float not above r*r, so a float compare is the same test.
    float r = t->ball_radius;
    if (r < 0.f) {
      if (t->kernel.type != MLOD_KERNEL_GAUSSIAN_MULTIPLE || t->kernel.n_sigmas < 1)
        return fail(MLOD_ERR_INVALID, "ball radius < 0 needs a multi-Gaussian kernel (3 sigmas[0])");
      r = t->kernel.sigmas[0] * 3;
    }
    const double r2 = (double)r * (double)r;
    float r2f = (float)r2;
    if ((double)r2f > r2) r2f = nextafterf(r2f, 0.f);
    k_ball_wave<FORM, KT><<<grid, NBS, 0, s>>>(t->view(), t->box, kp, t->qs.q, t->qs.qn, t->qs.perm, m, limit, nIter, r2f);
  }
  LAUNCH_CHECK();
  return MLOD_OK;
}

template <int FORM>
int launch_kt(mlod_tree* t, const KParams& kp, uint32_t limit, uint32_t nIter, hipStream_t s) {
  switch (kp.type) {
    case MLOD_KERNEL_GAUSSIAN: return launch_form<FORM, MLOD_KERNEL_GAUSSIAN>(t, kp, limit, nIter, s);
    case MLOD_KERNEL_SINGULAR: return launch_form<FORM, MLOD_KERNEL_SINGULAR>(t, kp, limit, nIter, s);
    default: return launch_form<FORM, MLOD_KERNEL_GAUSSIAN_MULTIPLE>(t, kp, limit, nIter, s);
  }
}

}  // namespace

int launch_neighbors(mlod_tree* t, uint32_t n, uint32_t n_iter, hipStream_t s) {
  const KParams kp = make_kparams(t->kernel);
  switch (t->precision) {
    case MLOD_PREC_FP32C: return launch_kt<1>(t, kp, n, n_iter, s);
    case MLOD_PREC_FP64: return launch_kt<2>(t, kp, n, n_iter, s);
    default: return launch_kt<0>(t, kp, n, n_iter, s);
  }
}

}  // namespace mlod

// project.hip — MLoD and global APSS projection kernels (gfx950).
//
// One lane per query (queries Morton-sorted by queries.hip, so a wavefront's 64 lanes walk
// neighbouring parts of the tree).  The MLoD sum of the reference,
//   value(leaf)  = sum over its points of w(p_i) stats_i                (Fastapss.cu:189-200)
//   value(far)   = w(pbar) stats                                        (Fastapss.cu:214-237)
//   value(near)  = sum_c [ b_c value(c) + (1 - b_c) share_c w(pbar) stats ] (Fastapss.cu:238-299,
//                  CPUUtils.cpp:478-501; b_c = (1 - cos(pi t))/2, t = dP / (dP + dC))
//   value(depth < minDepth) = sum_c value(c)                            (Fastapss.cu:303-329)
// is linear in the stats, so it is evaluated top-down: each node carries the product m of
// the blend factors on its path and adds m * (its own term) to one accumulator.  That needs
// no per-depth accumulator stack (the reference keeps 12 of them per thread); the per-lane
// traversal state is, per depth level, the first child, a mask of children still to descend,
// the node's multiplier and its distance term, held in LDS.  A near node loops over its
// contiguous children once: far children and leaves are summed right there, the parent's
// (1 - b_c) share is accumulated, and only near internal children are left in the mask.
//
// Arithmetic forms (template FORM):
//   0 plain FP32   — global-frame moments and the reference's FP32 sphere fit;
//   1 FP32C        — node-centred moments, contributions recentred on the query
//                    (sum w a (p - q), sum w a |p - q|^2, ...), Kahan-compensated FP32
//                    accumulation, FP64 sphere fit in the query frame;
//   2 plain FP64   — everything in double.
// Semantics (template SEM): 0 = REF_GPU (Fastapss.cu), 1 = REF_CPU (CPUUtils.cpp), see
// mlod.h.  Weight kernels (template KT): Kernel.h:37 numbering.
#include <cmath>

#include "mlod_internal.h"

namespace mlod {
namespace {

constexpr int PBS = 128;         // lanes per workgroup
// Levels of the LDS stack of waiting nodes.  A node waits only while it has more than one near
// child still to expand, and near children are internal nodes; internal nodes have depth
// 0..kMaxDepth - 1 and their children depth <= kMaxDepth, so a waiting node has depth at most
// kMaxDepth - 2 (its near children are internal, hence of depth <= kMaxDepth - 1): 11 levels,
// 13 B each per lane in the FP32 forms (18.3 KB per workgroup of 128 lanes).
constexpr int kLevels = kMaxDepth - 1;

#define LAUNCH_CHECK()                                                           \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) return hip_fail(e_, "kernel launch", __FILE__, __LINE__); \
  } while (0)

// ----------------------------------------------------------------------------- weights
// FP32 weights (Kernel.cpp:64-83 / GPUUtils.cu:243-257) from the squared distance.
// KT < 0: the kernel type is read at run time from the (wave-uniform) coefficients, so one
// launch serves workgroups of different kernel types with a scalar branch per evaluation.
template <int KT, class KP>
__device__ __forceinline__ float weight32(const KP& k, float d2) {
  if constexpr (KT < 0) {
    if (k.type == MLOD_KERNEL_GAUSSIAN_MULTIPLE) return weight32<MLOD_KERNEL_GAUSSIAN_MULTIPLE>(k, d2);
    if (k.type == MLOD_KERNEL_GAUSSIAN) return weight32<MLOD_KERNEL_GAUSSIAN>(k, d2);
    return weight32<MLOD_KERNEL_SINGULAR>(k, d2);
  } else if constexpr (KT == MLOD_KERNEL_GAUSSIAN_MULTIPLE) {
    // sum over s of sigma_s^-3 * 2^(a_s d2), in order of s
    float w = k.b[0] * __builtin_amdgcn_exp2f(d2 * k.a[0]);
    if (k.ns <= 8) {
      // statically indexed, so the coefficients sit in scalar registers; the wave-uniform exit
      // is one scalar compare and branch per term (the empty asm keeps the compiler from
      // hoisting the compares into lane masks), so no term beyond ns is evaluated
      int ns = k.ns;
#pragma unroll
      for (int s = 1; s < 8; ++s) {
        asm volatile("" : "+s"(ns));
        if (s >= ns) break;
        w = fmaf(k.b[s], __builtin_amdgcn_exp2f(d2 * k.a[s]), w);
      }
    } else {
      // more sigmas (the sample models' settings use up to 30): the first eight from scalar
      // registers, the rest four at a time, each group's coefficients fetched by two 16-byte
      // scalar loads (the table is 16-byte aligned) rather than one dependent load per term
#pragma unroll
      for (int s = 1; s < 8; ++s) w = fmaf(k.b[s], __builtin_amdgcn_exp2f(d2 * k.a[s]), w);
      int ns = k.ns;
      for (int s0 = 8; s0 < ns; s0 += 4) {
        const float4 A4 = *reinterpret_cast<const float4*>(k.a + s0);
        const float4 B4 = *reinterpret_cast<const float4*>(k.b + s0);
        w = fmaf(B4.x, __builtin_amdgcn_exp2f(d2 * A4.x), w);
        asm volatile("" : "+s"(ns));
        if (s0 + 1 >= ns) break;
        w = fmaf(B4.y, __builtin_amdgcn_exp2f(d2 * A4.y), w);
        asm volatile("" : "+s"(ns));
        if (s0 + 2 >= ns) break;
        w = fmaf(B4.z, __builtin_amdgcn_exp2f(d2 * A4.z), w);
        asm volatile("" : "+s"(ns));
        if (s0 + 3 >= ns) break;
        w = fmaf(B4.w, __builtin_amdgcn_exp2f(d2 * A4.w), w);
      }
    }
    return w == 0.f ? k.c : w;  // an all-underflowed sum falls back to constant_shift (Kernel.cpp:69-73)
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
  if constexpr (KT < 0) {
    if (k.type == MLOD_KERNEL_GAUSSIAN_MULTIPLE) return weight64<MLOD_KERNEL_GAUSSIAN_MULTIPLE>(k, d2);
    if (k.type == MLOD_KERNEL_GAUSSIAN) return weight64<MLOD_KERNEL_GAUSSIAN>(k, d2);
    return weight64<MLOD_KERNEL_SINGULAR>(k, d2);
  } else if constexpr (KT == MLOD_KERNEL_GAUSSIAN_MULTIPLE) {
    double w = 0.0;
    for (int s = 0; s < k.ns; ++s) w += exp(d2 * k.ad[s]) * k.bd[s];
    return w == 0.0 ? (double)k.c : w;
  } else if constexpr (KT == MLOD_KERNEL_GAUSSIAN) {
    return exp(d2 * k.ad[0]) + (double)k.c;
  } else {
    return pow((double)k.cp, (double)k.e) / pow(d2 + (double)k.c, (double)k.e);
  }
}

// Level blending (1 - cos(pi x)) / 2 (CPUUtils.cpp:400-408, Fastapss.cu:84-95).  The FP32
// form uses the hardware cosine, which takes revolutions: cos(pi x) = cos_rev(x / 2).
__device__ __forceinline__ float blend(float x) { return 0.5f - 0.5f * __builtin_amdgcn_cosf(0.5f * x); }
__device__ __forceinline__ double blend(double x) { return (1.0 - cos(3.14159265 * x)) * 0.5; }

__device__ __forceinline__ float fsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }

// ----------------------------------------------------------------------------- accumulators
struct KSum {
  float s, c;
  __device__ __forceinline__ void init() { s = 0.f, c = 0.f; }
  __device__ __forceinline__ void add(float x) {
    const float y = __fsub_rn(x, c);
    const float t = __fadd_rn(s, y);
    c = __fsub_rn(__fsub_rn(t, s), y);
    s = t;
  }
  __device__ __forceinline__ double val() const { return (double)s - (double)c; }
};

template <class T>
struct V3 {
  T x, y, z;
};

// FORM 0: plain global-frame FP32 sums (apssStatsGPU, GPUUtils.cu:318-392)
struct AccP32 {
  float w, px, py, pz, nx, ny, nz, pn, pp;
  __device__ __forceinline__ void init() { w = px = py = pz = nx = ny = nz = pn = pp = 0.f; }
};
// FORM 1: compensated, query-frame.  Terms go by one FMA each into plain FP32 partial sums (the
// l* members), which are added into the Kahan-compensated sums every kFlush traversal iterations
// and at the end of a walk (flush()): the compensation is paid once per block of terms, and a
// block's rounding error is bounded by its few terms.
constexpr uint32_t kFlush = 8;
struct AccC32 {
  KSum w, px, py, pz, nx, ny, nz, pn, pp;
  float lw, lpx, lpy, lpz, lnx, lny, lnz, lpn, lpp;
  __device__ __forceinline__ void init() {
    w.init(), px.init(), py.init(), pz.init(), nx.init(), ny.init(), nz.init(), pn.init(), pp.init();
    lw = lpx = lpy = lpz = lnx = lny = lnz = lpn = lpp = 0.f;
  }
  __device__ __forceinline__ void flush() {
    w.add(lw), px.add(lpx), py.add(lpy), pz.add(lpz), nx.add(lnx), ny.add(lny), nz.add(lnz), pn.add(lpn), pp.add(lpp);
    lw = lpx = lpy = lpz = lnx = lny = lnz = lpn = lpp = 0.f;
  }
};
// FORM 2: plain FP64
struct AccP64 {
  double w, px, py, pz, nx, ny, nz, pn, pp;
  __device__ __forceinline__ void init() { w = px = py = pz = nx = ny = nz = pn = pp = 0.0; }
};

template <int FORM> struct FormT;
template <> struct FormT<0> { using T = float;  using Acc = AccP32; };
template <> struct FormT<1> { using T = float;  using Acc = AccC32; };
template <> struct FormT<2> { using T = double; using Acc = AccP64; };

struct Counters {
  uint32_t far_, leafPts, tested, nearN, iters;
};

// Contribution of one point (leaf member) with multiplier f, weight at the point's mean
// position (a p) / a (Fastapss.cu:193; CPUUtils.cpp:455).
template <int FORM, int KT, class KP>
__device__ __forceinline__ void add_point(typename FormT<FORM>::Acc& acc, const KP& kp, const TreeView& tv, uint32_t j,
                                          V3<typename FormT<FORM>::T> q, typename FormT<FORM>::T f) {
  const float4 P = tv.pts[j], N = tv.nrm[j];
  if constexpr (FORM == 0) {
    const float a = P.w;
    const float apx = a * P.x, apy = a * P.y, apz = a * P.z;
    const float dx = apx / a - q.x, dy = apy / a - q.y, dz = apz / a - q.z;
    const float w = f * weight32<KT>(kp, dx * dx + dy * dy + dz * dz);
    acc.w = fmaf(w, a, acc.w);
    acc.px = fmaf(w, apx, acc.px);
    acc.py = fmaf(w, apy, acc.py);
    acc.pz = fmaf(w, apz, acc.pz);
    acc.nx = fmaf(w, a * N.x, acc.nx);
    acc.ny = fmaf(w, a * N.y, acc.ny);
    acc.nz = fmaf(w, a * N.z, acc.nz);
    acc.pn = fmaf(w, a * (P.x * N.x + P.y * N.y + P.z * N.z), acc.pn);
    acc.pp = fmaf(w, a * (P.x * P.x + P.y * P.y + P.z * P.z), acc.pp);
  } else if constexpr (FORM == 1) {
    const float dx = P.x - q.x, dy = P.y - q.y, dz = P.z - q.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float fa = f * weight32<KT>(kp, d2) * P.w;
    acc.lw += fa;
    acc.lpx = fmaf(fa, dx, acc.lpx);
    acc.lpy = fmaf(fa, dy, acc.lpy);
    acc.lpz = fmaf(fa, dz, acc.lpz);
    acc.lnx = fmaf(fa, N.x, acc.lnx);
    acc.lny = fmaf(fa, N.y, acc.lny);
    acc.lnz = fmaf(fa, N.z, acc.lnz);
    acc.lpn = fmaf(fa, dx * N.x + dy * N.y + dz * N.z, acc.lpn);
    acc.lpp = fmaf(fa, d2, acc.lpp);
  } else {
    const double a = P.w, x = P.x, y = P.y, z = P.z;
    const double apx = a * x, apy = a * y, apz = a * z;
    const double dx = apx / a - q.x, dy = apy / a - q.y, dz = apz / a - q.z;
    const double w = f * weight64<KT>(kp, dx * dx + dy * dy + dz * dz);
    acc.w += w * a;
    acc.px += w * apx;
    acc.py += w * apy;
    acc.pz += w * apz;
    acc.nx += w * (a * N.x);
    acc.ny += w * (a * N.y);
    acc.nz += w * (a * N.z);
    acc.pn += w * (a * ((x * N.x + y * N.y) + z * N.z));
    acc.pp += w * (a * ((x * x + y * y) + z * z));
  }
}

// Moments of one node as loaded from HBM (node-centred FP32 for the compensated form,
// global-frame FP32 for the plain form, FP64 for the FP64 form).
template <int FORM> struct NodeMom { float4 m0, m1; };
template <> struct NodeMom<2> { double4 m0, m1; double A; };

// Record `idx` of `bytes` bytes from `base`, with the byte offset formed in 32 bits (the tree
// build guarantees node counts below 2^27), so the address is a scalar base plus one 32-bit
// vector offset instead of 64-bit vector arithmetic.
template <class T>
__device__ __forceinline__ T ld_rec(const T* base, uint32_t idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (idx * (uint32_t)sizeof(T)));
}

template <int FORM>
__device__ __forceinline__ NodeMom<FORM> load_mom(const TreeView& tv, uint32_t c) {
  NodeMom<FORM> r;
  if constexpr (FORM == 0) {
    r.m0 = ld_rec(tv.momP, 2 * c), r.m1 = ld_rec(tv.momP, 2 * c + 1);
  } else if constexpr (FORM == 1) {
    r.m0 = ld_rec(tv.momC, 2 * c), r.m1 = ld_rec(tv.momC, 2 * c + 1);
  } else {
    r.m0 = tv.momD[2 * (size_t)c], r.m1 = tv.momD[2 * (size_t)c + 1];
    r.A = tv.AD[c];
  }
  return r;
}

// Contribution of a node's moments with multiplier f, weight at its mean point.
template <int FORM, int KT, class KP>
__device__ __forceinline__ void add_mom(typename FormT<FORM>::Acc& acc, const KP& kp, const NodeMom<FORM>& mm, float Af,
                                        V3<typename FormT<FORM>::T> q, typename FormT<FORM>::T f) {
  if constexpr (FORM == 0) {
    const float4 m0 = mm.m0, m1 = mm.m1;
    const float dx = m0.x / Af - q.x, dy = m0.y / Af - q.y, dz = m0.z / Af - q.z;
    const float w = f * weight32<KT>(kp, dx * dx + dy * dy + dz * dz);
    acc.w = fmaf(w, Af, acc.w);
    acc.px = fmaf(w, m0.x, acc.px);
    acc.py = fmaf(w, m0.y, acc.py);
    acc.pz = fmaf(w, m0.z, acc.pz);
    acc.nx = fmaf(w, m1.x, acc.nx);
    acc.ny = fmaf(w, m1.y, acc.ny);
    acc.nz = fmaf(w, m1.z, acc.nz);
    acc.pn = fmaf(w, m0.w, acc.pn);
    acc.pp = fmaf(w, m1.w, acc.pp);
  } else if constexpr (FORM == 1) {
    const float4 m0 = mm.m0, m1 = mm.m1;
    const float dx = m0.x - q.x, dy = m0.y - q.y, dz = m0.z - q.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float fw = f * weight32<KT>(kp, d2);
    const float fa = fw * Af;
    acc.lw += fa;
    acc.lpx = fmaf(fa, dx, acc.lpx);
    acc.lpy = fmaf(fa, dy, acc.lpy);
    acc.lpz = fmaf(fa, dz, acc.lpz);
    acc.lnx = fmaf(fw, m1.x, acc.lnx);
    acc.lny = fmaf(fw, m1.y, acc.lny);
    acc.lnz = fmaf(fw, m1.z, acc.lnz);
    acc.lpn = fmaf(fw, m0.w + (dx * m1.x + dy * m1.y + dz * m1.z), acc.lpn);
    acc.lpp = fmaf(fw, m1.w + Af * d2, acc.lpp);
  } else {
    const double4 m0 = mm.m0, m1 = mm.m1;
    const double A = mm.A;
    const double dx = m0.x / A - q.x, dy = m0.y / A - q.y, dz = m0.z / A - q.z;
    const double w = f * weight64<KT>(kp, dx * dx + dy * dy + dz * dz);
    acc.w += w * A;
    acc.px += w * m0.x;
    acc.py += w * m0.y;
    acc.pz += w * m0.z;
    acc.nx += w * m1.x;
    acc.ny += w * m1.y;
    acc.nz += w * m1.z;
    acc.pn += w * m0.w;
    acc.pp += w * m1.w;
  }
}

template <int FORM, int KT, class KP>
__device__ __forceinline__ void add_node(typename FormT<FORM>::Acc& acc, const KP& kp, const TreeView& tv, uint32_t c,
                                         float Af, V3<typename FormT<FORM>::T> q, typename FormT<FORM>::T f) {
  add_mom<FORM, KT>(acc, kp, load_mom<FORM>(tv, c), Af, q, f);
}

template <int FORM, int SEM, int KT, bool CNT, class KP>
__device__ __forceinline__ void add_leaf(typename FormT<FORM>::Acc& acc, const KP& kp, const TreeView& tv, uint4 h,
                                         V3<typename FormT<FORM>::T> q, typename FormT<FORM>::T f, Counters& ct) {
  const uint32_t first = h.x, cnt = hdr_count(h.y);
  if constexpr (SEM == MLOD_SEM_REF_CPU) {
    // apssStats of a leaf is assigned, not summed, per point: only the last point stays
    // (CPUUtils.cpp:448-457).
    add_point<FORM, KT>(acc, kp, tv, first + cnt - 1, q, f);
    if constexpr (CNT) ct.leafPts += 1;
  } else {
    for (uint32_t j = first; j < first + cnt; ++j) add_point<FORM, KT>(acc, kp, tv, j, q, f);
    if constexpr (CNT) ct.leafPts += cnt;
  }
}

template <class T>
__device__ __forceinline__ T dist3(V3<T> q, float4 g) {
  const T dx = q.x - (T)g.x, dy = q.y - (T)g.y, dz = q.z - (T)g.z;
  return fsqrt(dx * dx + dy * dy + dz * dz);
}

// Traversal arithmetic per form.  The plain forms keep IEEE division and square root (the
// reference's arithmetic); the compensated form, whose error budget is set by its FP64 fit and
// compensated sums, takes the hardware reciprocal and square root (1 ulp).
template <int FORM>
__device__ __forceinline__ typename FormT<FORM>::T tdist(V3<typename FormT<FORM>::T> q, float4 g) {
  if constexpr (FORM == 1) {
    const float dx = q.x - g.x, dy = q.y - g.y, dz = q.z - g.z;
    return __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
  } else {
    return dist3(q, g);
  }
}
template <int FORM, class T>
__device__ __forceinline__ T tratio(T a, T b) {  // a / b
  if constexpr (FORM == 1)
    return a * __builtin_amdgcn_rcpf(b);
  else
    return a / b;
}

// A node whose near children wait, per level and lane: what its next near child needs to start
// (the node's first child, multiplier and distance term), so a lane resumes it without reloading it.
template <class T>
struct LevelState {
  uint32_t (*first)[PBS];  // the node's first child
  uint8_t (*mask)[PBS];    // its near children still to expand
  T (*mul)[PBS];           // its multiplier
  T (*dist)[PBS];          // its distance term |d - cutoff r| (negative: no blending below it)
};

// Flattened traversal.  The lane's current node (expanded: near, or above minDepth) has its
// children tested one per loop iteration; when they are done the node's own (1 - b_c) share
// is added and its first near child becomes the current node.  Either way an iteration handles
// exactly one node X: it loads X, computes its distance and its blend factor against its
// parent, then either classifies X as a child (far / leaf / near) or makes it the current node,
// and ends in one shared node contribution.  Lanes of a wavefront therefore advance by one node
// per iteration whatever their position in their own tree walk, instead of waiting for the
// widest child loop of the wavefront at every level.
//
// Each iteration loads X's geometry and header, and the moments of the node it contributes (X
// itself when testing, the finished node when moving on).  A finished node with more than one
// near child leaves {first child, remaining near children, multiplier, distance term} in LDS for
// its depth (13 B per level and lane, a per-lane bit set marks the levels in use); when a lane's
// path runs out, the deepest such node's next near child is taken from there directly.
//
// The walk is resumable: begin() sets it up for one query step, iter() advances it by one
// iteration, and `active` drops when the step's sum is complete, so a kernel can interleave
// the walks of successive queries on one lane.
template <int FORM, int SEM, int KT, bool CNT, class KP>
struct Walker {
  using T = typename FormT<FORM>::T;
  using Acc = typename FormT<FORM>::Acc;
  uint32_t node, first, nc, k, mask, pend;
  int lvl;
  bool blendNode, active;
  T m, distP, shareSum, shareScale;
  float Ai;

  __device__ __forceinline__ void begin(Acc& acc, const KP& kp, const TreeView& tv, V3<T> q, T cutoff,
                                        uint32_t minDepth, Counters& ct) {
    active = false;
    const uint4 h0 = tv.hdr[0];
    if (hdr_leaf(h0.y)) {
      add_leaf<FORM, SEM, KT, CNT>(acc, kp, tv, h0, q, T(1), ct);
      return;
    }
    const float4 g0 = tv.geo[0];
    const T d0 = tdist<FORM>(q, g0);
    if (minDepth == 0 && d0 > (T)g0.w * cutoff) {
      add_node<FORM, KT>(acc, kp, tv, 0, __uint_as_float(h0.z), q, T(1));
      if constexpr (CNT) ct.far_ += 1;
      return;
    }
    // current node = root
    node = 0, first = h0.x, nc = hdr_count(h0.y), k = 0, mask = 0, pend = 0;
    lvl = 0;
    blendNode = minDepth == 0;
    m = T(1);
    distP = blendNode ? fabs(d0 - (T)g0.w * cutoff) : T(-1);
    shareSum = T(0);
    Ai = __uint_as_float(h0.z);
    shareScale = SEM == MLOD_SEM_REF_CPU ? tratio<FORM>(T(1), (T)nc) : tratio<FORM>(T(1), (T)Ai);
    active = true;
  }

  __device__ __forceinline__ void iter(Acc& acc, const KP& kp, const TreeView& tv, V3<T> q, T cutoff,
                                       uint32_t minDepth, LevelState<T> st, int tid, Counters& ct) {
    if constexpr (CNT) ct.iters += 1;
    // --- pick X: the next child of the current node, or the next near node to expand (of the
    // finished node, or of the deepest node whose near children wait)
    const bool testing = k < nc;
    uint32_t x = 0;
    T mp = T(0), dp = T(-1);
    bool ownContrib = false;
    const uint32_t ownNode = node;
    const float ownA = Ai;
    T ownF = T(0);
    if (testing) {
      x = first + k;
      mp = m;
      dp = blendNode ? distP : T(-1);
    } else {
      ownContrib = blendNode && shareSum != T(0);
      ownF = m * shareSum;
      if constexpr (CNT) ct.nearN += (blendNode && shareSum != T(0)) ? 1u : 0u;
      if (mask) {
        // the finished node's first near child is next; the rest wait at this level
        const uint32_t rest = mask & (mask - 1u);
        mp = m;
        dp = blendNode ? distP : T(-1);
        if (rest) {
          st.first[lvl][tid] = first;
          st.mask[lvl][tid] = (uint8_t)rest;
          st.mul[lvl][tid] = mp;
          st.dist[lvl][tid] = dp;
          pend |= 1u << lvl;
        }
        x = first + (uint32_t)__builtin_ctz(mask);
        lvl = lvl + 1;
      } else if (pend) {
        // the deepest waiting node's next near child
        const int L = 31 - __builtin_clz(pend);
        const uint32_t wm = st.mask[L][tid];
        const uint32_t rest = wm & (wm - 1u);
        x = st.first[L][tid] + (uint32_t)__builtin_ctz(wm);
        mp = st.mul[L][tid];
        dp = st.dist[L][tid];
        if (rest)
          st.mask[L][tid] = (uint8_t)rest;
        else
          pend &= ~(1u << L);
        lvl = L + 1;
      } else {
        active = false;
      }
    }
    // --- loads: X's geometry and header, and the moments that may be added
    const uint32_t tgt = testing ? x : ownNode;
    const NodeMom<FORM> mm = load_mom<FORM>(tv, tgt);
    float4 gx = make_float4(0.f, 0.f, 0.f, 0.f);
    uint4 hx = make_uint4(0u, 0u, 0u, 0u);
    if (active) {
      gx = ld_rec(tv.geo, x);
      hx = ld_rec(tv.hdr, x);
    }
    // --- X's blend factor against its parent (CPUUtils.cpp:489-497)
    bool contrib = ownContrib;
    float At = ownA;
    T f = ownF;
    if (active) {
      const T dxq = tdist<FORM>(q, gx);
      const T rc = (T)gx.w * cutoff;
      T beta = T(1);
      if (dp >= T(0)) {
        const T dc = fmax(dxq - rc, T(0));
        beta = blend(tratio<FORM>(dp, dp + dc));
      }
      const T mx = mp * beta;
      const float Ax = __uint_as_float(hx.z);
      if (testing) {
        ++k;
        if (blendNode) {
          const T share = SEM == MLOD_SEM_REF_CPU ? shareScale
                                                  : (FORM == 1 ? (T)Ax * shareScale : (T)Ax / (T)Ai);
          shareSum += (T(1) - beta) * share;
        }
        if constexpr (CNT) ct.tested += 1;
        const bool leaf = hdr_leaf(hx.y);
        const bool single = leaf && hdr_count(hx.y) == 1u;
        const bool far = !leaf && hdr_depth(hx.y) >= minDepth && dxq > rc;
        if (leaf && !single) add_leaf<FORM, SEM, KT, CNT>(acc, kp, tv, hx, q, mx, ct);  // multi-point leaf
        if (!leaf && !far) mask |= 1u << (k - 1u);
        if constexpr (CNT) {
          ct.far_ += far ? 1u : 0u;
          ct.leafPts += single ? 1u : 0u;
        }
        contrib = far || single;
        At = Ax;
        f = mx;
      } else {
        // X becomes the current node, its children to test
        node = x;
        first = hx.x;
        nc = hdr_count(hx.y);
        k = 0u;
        mask = 0u;
        blendNode = hdr_depth(hx.y) >= minDepth;
        distP = fabs(dxq - rc);
        m = mx;
        shareSum = T(0);
        Ai = Ax;
        shareScale = SEM == MLOD_SEM_REF_CPU ? tratio<FORM>(T(1), (T)nc) : tratio<FORM>(T(1), (T)Ax);
      }
    }
    if (contrib) add_mom<FORM, KT>(acc, kp, mm, At, q, f);
  }
};

// The whole walk of one query step on one lane.
template <int FORM, int SEM, int KT, bool CNT, class KP>
__device__ __forceinline__ void traverse(typename FormT<FORM>::Acc& acc, const KP& kp, const TreeView& tv,
                                         V3<typename FormT<FORM>::T> q, typename FormT<FORM>::T cutoff,
                                         uint32_t minDepth, LevelState<typename FormT<FORM>::T> st, int tid,
                                         Counters& ct) {
  Walker<FORM, SEM, KT, CNT, KP> w;
  w.begin(acc, kp, tv, q, cutoff, minDepth, ct);
  // the lanes of a wavefront start together and advance one iteration per loop, so the
  // compensated form's block flushes fall on the same loop iterations for all of them
  uint32_t itc = 0;
  while (w.active) {
    w.iter(acc, kp, tv, q, cutoff, minDepth, st, tid, ct);
    if constexpr (FORM == 1) {
      if (++itc == kFlush) {
        acc.flush();
        itc = 0;
      }
    }
  }
  if constexpr (FORM == 1) acc.flush();
}

// ----------------------------------------------------------------------------- wave-shared walk
// The same MLoD sum, walked once per wavefront over the union of its lanes' walks.  The node
// being handled is wave-uniform, so its records come through scalar loads; each lane keeps
// its own multiplier, distance term, share sum and near-child bits and takes part only in the
// nodes its own walk visits (a lane is active at a child when it was active at the parent and
// found the child near).  The union walk is a depth-first walk in child order, like each lane's
// own walk, so every lane adds the same terms in the same order as walking alone, and it counts
// its own iterations (one per tested child and one per finished node), so the compensated
// form's block flushes fall where they fall in a lane's own walk.
struct WaveLevel {
  uint32_t first;  // the waiting node's first child
  uint32_t rest;   // its children still to expand (near for at least one lane)
  uint64_t act;    // the lanes active at the waiting node
};

template <class T>
struct WaveStack {
  WaveLevel (*lv)[PBS / 64];  // per level and wavefront
  T (*mul)[PBS];              // per level and lane: the waiting node's multiplier
  T (*dist)[PBS];             // its distance term (negative: no blending below it)
  uint8_t (*near)[PBS];       // its near children for this lane
};

// A wave-uniform record through the constant address space: the tree is not written while a
// projection runs, so the load is a scalar load into SGPRs (one per wavefront, through the
// scalar cache) instead of a vector load per lane.
template <class T>
__device__ __forceinline__ T ld_uni(const T* base, uint32_t idx) {
  static_assert(sizeof(T) % 4 == 0, "records are whole dwords");
  typedef uint32_t W __attribute__((ext_vector_type(sizeof(T) / 4)));
  const W w = *(const __attribute__((address_space(4))) W*)((const __attribute__((address_space(4))) char*)base +
                                                            (uint64_t)idx * sizeof(T));
  T r;
  __builtin_memcpy(&r, &w, sizeof(T));
  return r;
}

template <int FORM>
__device__ __forceinline__ NodeMom<FORM> load_mom_uni(const TreeView& tv, uint32_t c) {
  NodeMom<FORM> r;
  if constexpr (FORM == 0) {
    r.m0 = ld_uni(tv.momP, 2 * c), r.m1 = ld_uni(tv.momP, 2 * c + 1);
  } else if constexpr (FORM == 1) {
    r.m0 = ld_uni(tv.momC, 2 * c), r.m1 = ld_uni(tv.momC, 2 * c + 1);
  } else {
    r.m0 = ld_uni(tv.momD, 2 * c), r.m1 = ld_uni(tv.momD, 2 * c + 1);
    r.A = ld_uni(tv.AD, c);
  }
  return r;
}

template <int FORM, int SEM, int KT, bool CNT>
__device__ __forceinline__ void traverse_wave(typename FormT<FORM>::Acc& acc, const KParams& kp, const TreeView& tv,
                                              V3<typename FormT<FORM>::T> q, typename FormT<FORM>::T cutoff,
                                              uint32_t minDepth, bool live, WaveStack<typename FormT<FORM>::T> st,
                                              int tid, int wv, Counters& ct, uint32_t& waveIters) {
  using T = typename FormT<FORM>::T;
  const uint32_t lane = __lane_id();
  uint32_t itc = 0;  // the lane's own iterations since the last block flush (compensated form)
  const uint4 h0 = ld_uni(tv.hdr, 0u);
  if (hdr_leaf(h0.y)) {
    if (live) add_leaf<FORM, SEM, KT, CNT>(acc, kp, tv, h0, q, T(1), ct);
    return;
  }
  const float4 g0 = ld_uni(tv.geo, 0u);
  bool blendNode = minDepth == 0;
  bool act = live;
  T m = T(1), distP = T(-1), shareSum = T(0);
  uint32_t nearb = 0;
  if (act) {
    const T d0 = tdist<FORM>(q, g0);
    if (minDepth == 0 && d0 > (T)g0.w * cutoff) {
      add_mom<FORM, KT>(acc, kp, load_mom_uni<FORM>(tv, 0u), __uint_as_float(h0.z), q, T(1));
      if constexpr (CNT) ct.far_ += 1;
      act = false;
    } else {
      distP = blendNode ? fabs(d0 - (T)g0.w * cutoff) : T(-1);
    }
  }
  uint64_t actMask = __ballot(act);
  if (!actMask) return;
  // the current node X (wave-uniform)
  uint32_t first = h0.x, nc = hdr_count(h0.y), k = 0, un = 0, pend = 0;
  int lvl = 0;
  float Ai = __uint_as_float(h0.z);
  T shareScale = SEM == MLOD_SEM_REF_CPU ? tratio<FORM>(T(1), (T)nc) : tratio<FORM>(T(1), (T)Ai);
  NodeMom<FORM> momX = load_mom_uni<FORM>(tv, 0);
  while (true) {
    if constexpr (CNT) ++waveIters;
    // one iteration: test child k of X, or finish X (its own (1 - b_c) share); either way at
    // most one node term per lane, added at the one call site below
    const bool testing = k < nc;
    NodeMom<FORM> mm = momX;
    float At = Ai;
    T f = m * shareSum;
    bool contrib = false;
    if (testing) {
      const uint32_t x = __builtin_amdgcn_readfirstlane(first + k);
      const float4 gx = ld_uni(tv.geo, x);
      const uint4 hx = ld_uni(tv.hdr, x);
      mm = load_mom_uni<FORM>(tv, x);
      const float Ax = __uint_as_float(hx.z);
      const bool leaf = hdr_leaf(hx.y);
      const bool single = leaf && hdr_count(hx.y) == 1u;
      const bool deep = hdr_depth(hx.y) >= minDepth;
      bool nearc = false;
      if (act) {
        if constexpr (CNT) ct.iters += 1, ct.tested += 1;
        const T dxq = tdist<FORM>(q, gx);
        const T rc = (T)gx.w * cutoff;
        const T dp = blendNode ? distP : T(-1);
        T beta = T(1);
        if (dp >= T(0)) {
          const T dc = fmax(dxq - rc, T(0));
          beta = blend(tratio<FORM>(dp, dp + dc));
        }
        const T mx = m * beta;
        if (blendNode) {
          const T share = SEM == MLOD_SEM_REF_CPU ? shareScale : (FORM == 1 ? (T)Ax * shareScale : (T)Ax / (T)Ai);
          shareSum += (T(1) - beta) * share;
        }
        const bool far = !leaf && deep && dxq > rc;
        if (leaf && !single) add_leaf<FORM, SEM, KT, CNT>(acc, kp, tv, hx, q, mx, ct);
        nearc = !leaf && !far;
        if (nearc) nearb |= 1u << k;
        if constexpr (CNT) {
          ct.far_ += far ? 1u : 0u;
          ct.leafPts += single ? 1u : 0u;
        }
        contrib = far || single;
        f = mx;
      }
      At = Ax;
      if (__ballot(nearc)) un |= 1u << k;
      ++k;
    } else if (act) {
      if constexpr (CNT) {
        ct.iters += 1;
        ct.nearN += (blendNode && shareSum != T(0)) ? 1u : 0u;
      }
      contrib = blendNode && shareSum != T(0);
    }
    if (contrib) add_mom<FORM, KT>(acc, kp, mm, At, q, f);
    if constexpr (FORM == 1) {
      if (act && ++itc == kFlush) {
        acc.flush();
        itc = 0;
      }
    }
    if (testing) continue;
    // X is finished: the next node to expand
    uint32_t firstP, c, nbits;
    uint64_t actP;
    T mp, dp;
    if (un) {
      // X's first near child is next; the rest wait at this level
      c = __builtin_ctz(un);
      const uint32_t rest = un & (un - 1u);
      mp = m;
      dp = blendNode ? distP : T(-1);
      nbits = nearb;
      actP = actMask;
      firstP = first;
      if (rest) {
        if (lane == 0) st.lv[lvl][wv] = WaveLevel{first, rest, actMask};
        st.mul[lvl][tid] = mp;
        st.dist[lvl][tid] = dp;
        st.near[lvl][tid] = (uint8_t)nearb;
        pend |= 1u << lvl;
      }
      lvl = lvl + 1;
    } else if (pend) {
      // the deepest waiting node's next near child
      const int L = 31 - __builtin_clz(pend);
      const WaveLevel e = st.lv[L][wv];
      firstP = __builtin_amdgcn_readfirstlane(e.first);
      const uint32_t wm = __builtin_amdgcn_readfirstlane(e.rest);
      actP = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(e.act >> 32)) << 32) |
             (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)e.act);
      c = __builtin_ctz(wm);
      const uint32_t rest = wm & (wm - 1u);
      if (rest) {
        if (lane == 0) st.lv[L][wv].rest = rest;
      } else {
        pend &= ~(1u << L);
      }
      mp = st.mul[L][tid];
      dp = st.dist[L][tid];
      nbits = st.near[L][tid];
      lvl = L + 1;
    } else {
      break;
    }
    // descend: the child becomes X for the lanes that found it near
    act = ((actP >> lane) & 1u) && ((nbits >> c) & 1u);
    actMask = __ballot(act);
    const uint32_t xn = __builtin_amdgcn_readfirstlane(firstP + c);
    const float4 gx = ld_uni(tv.geo, xn);
    const uint4 hx = ld_uni(tv.hdr, xn);
    momX = load_mom_uni<FORM>(tv, xn);
    if (act) {
      const T dxq = tdist<FORM>(q, gx);
      const T rc = (T)gx.w * cutoff;
      T beta = T(1);
      if (dp >= T(0)) {
        const T dc = fmax(dxq - rc, T(0));
        beta = blend(tratio<FORM>(dp, dp + dc));
      }
      m = mp * beta;
      distP = fabs(dxq - rc);
    }
    first = hx.x;
    nc = hdr_count(hx.y);
    k = 0;
    un = 0;
    nearb = 0;
    shareSum = T(0);
    blendNode = hdr_depth(hx.y) >= minDepth;
    Ai = __uint_as_float(hx.z);
    shareScale = SEM == MLOD_SEM_REF_CPU ? tratio<FORM>(T(1), (T)nc) : tratio<FORM>(T(1), (T)Ai);
  }
}

// ----------------------------------------------------------------------------- sphere fit
// Algebraic sphere fit and projection, Fastapss.cu:354-395 (REF_GPU: plane below |u4| 1e-6,
// normalize(0) = 0) and CPUUtils.cpp:739-782 (REF_CPU: plane below 1e-12, normalize(0) = NaN).
template <class T, int SEM>
__device__ __forceinline__ V3<T> nrmz(V3<T> v) {
  const T l2 = v.x * v.x + v.y * v.y + v.z * v.z;
  if (SEM == MLOD_SEM_REF_GPU && l2 == T(0)) return V3<T>{0, 0, 0};
  const T inv = T(1) / fsqrt(l2);
  return V3<T>{v.x * inv, v.y * inv, v.z * inv};
}

template <class T>
__device__ __forceinline__ T dot3(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Sums in frame `o` (origin of the accumulated positions): q is the query in that frame.
// Returns the projection in that frame (NaN guard and step clamp applied).
// STABLE (query frame only, q = 0): the sphere point c + r (q - c)/|q - c| is computed as
// tau * u/|u| with tau = -2 u0 / (|u| + sqrt(|u|^2 - 4 u4 u0)), the small root of
// u0 + tau |u| + u4 tau^2 = 0 along the line through q and the centre.  It is the same point
// (|c| - r = 2 u0 / (|u| + sqrt(|u|^2 - 4 u4 u0)) up to sign) without the cancellation of c and
// r near planar fits (|u4| -> 0, |c| ~ 1/|u4|), and tends to the plane projection as u4 -> 0.
// kind: 0 sphere, 1 plane (|u4| below the threshold), 2 invalid (NaN: the query is kept)
template <class T, class TF, int SEM, bool STABLE = false>
__device__ __forceinline__ void fit(T sw, V3<T> sp, V3<T> sn, T spn, T spp, V3<T> q, T stepMax, V3<T>& outP, V3<T>& outN,
                                    float& kind) {
  const T A = spn - dot3(sp, V3<T>{sn.x / sw, sn.y / sw, sn.z / sw});
  const T B = spp - dot3(sp, V3<T>{sp.x / sw, sp.y / sw, sp.z / sw});
  const T u4 = (T)((0.5 * (double)A) / (double)B);
  const T two_u4 = T(2) * u4;
  const V3<T> u{(sn.x - two_u4 * sp.x) / sw, (sn.y - two_u4 * sp.y) / sw, (sn.z - two_u4 * sp.z) / sw};
  const T u0 = -(dot3(sp, u) + u4 * spp) / sw;
  V3<T> p = q, nr;
  const double tau = SEM == MLOD_SEM_REF_GPU ? 0.000001 : 0.000000000001;
  kind = fabs((double)u4) < tau ? 1.f : 0.f;
  if (fabs((double)u4) < tau) {
    const V3<T> nn{-u.x, -u.y, -u.z};
    const T lambda = (u0 - dot3(p, nn)) / dot3(nn, nn);
    p = V3<T>{p.x + lambda * nn.x, p.y + lambda * nn.y, p.z + lambda * nn.z};
    nr = nrmz<T, SEM>(sn);
  } else if constexpr (STABLE) {
    const T un2 = dot3(u, u);
    const T un = fsqrt(un2);
    const T disc = un2 - T(4) * u4 * u0;
    if (disc < T(0)) {  // r^2 < 0: the reference clamps r to 0, i.e. the centre
      const T m2 = T(-2) * u4;
      p = V3<T>{u.x / m2, u.y / m2, u.z / m2};
    } else {
      const T s = (T(-2) * u0 / (un + fsqrt(disc))) / un;
      p = V3<T>{s * u.x, s * u.y, s * u.z};
    }
    nr = nrmz<T, SEM>(V3<T>{u.x + two_u4 * p.x, u.y + two_u4 * p.y, u.z + two_u4 * p.z});
  } else {
    const T m2 = T(-2) * u4;
    const V3<T> c{u.x / m2, u.y / m2, u.z / m2};
    const T v = dot3(c, c) - u0 / u4;
    const T r = fsqrt(v > T(0) ? v : T(0));
    const V3<T> pc = nrmz<T, SEM>(V3<T>{p.x - c.x, p.y - c.y, p.z - c.z});
    p = V3<T>{c.x + r * pc.x, c.y + r * pc.y, c.z + r * pc.z};
    nr = nrmz<T, SEM>(V3<T>{u.x + two_u4 * p.x, u.y + two_u4 * p.y, u.z + two_u4 * p.z});
  }
  if (isnan(p.x) || isnan(p.y) || isnan(p.z)) {
    outP = q;
    outN = V3<T>{0, 0, 0};
    kind = 2.f;
    return;
  }
  const V3<T> dlt{p.x - q.x, p.y - q.y, p.z - q.z};
  const T step = fsqrt(dot3(dlt, dlt));
  if (step > stepMax) {
    const T s = stepMax / step;
    p = V3<T>{q.x + s * dlt.x, q.y + s * dlt.y, q.z + s * dlt.z};
  }
  outP = p;
  outN = nr;
}

// One projection step of query q (double carried between steps).
template <int FORM, int SEM>
__device__ __forceinline__ void fit_step(const typename FormT<FORM>::Acc& acc, V3<double>& qd, V3<float> qf, float stepMax,
                                         float4& nout) {
  if constexpr (FORM == 0) {
    V3<float> p, n;
    float kind;
    fit<float, float, SEM>(acc.w, V3<float>{acc.px, acc.py, acc.pz}, V3<float>{acc.nx, acc.ny, acc.nz}, acc.pn, acc.pp, qf,
                           stepMax, p, n, kind);
    qd = V3<double>{p.x, p.y, p.z};
    nout = make_float4(n.x, n.y, n.z, kind);
  } else if constexpr (FORM == 1) {
    // query frame: the query is the origin
    V3<double> p, n;
    float kind;
    fit<double, double, SEM, true>(acc.w.val(), V3<double>{acc.px.val(), acc.py.val(), acc.pz.val()},
                             V3<double>{acc.nx.val(), acc.ny.val(), acc.nz.val()}, acc.pn.val(), acc.pp.val(),
                             V3<double>{0, 0, 0}, (double)stepMax, p, n, kind);
    qd = V3<double>{(double)qf.x + p.x, (double)qf.y + p.y, (double)qf.z + p.z};
    nout = make_float4((float)n.x, (float)n.y, (float)n.z, kind);
  } else {
    V3<double> p, n;
    float kind;
    fit<double, double, SEM>(acc.w, V3<double>{acc.px, acc.py, acc.pz}, V3<double>{acc.nx, acc.ny, acc.nz}, acc.pn, acc.pp,
                             qd, (double)stepMax, p, n, kind);
    qd = p;
    nout = make_float4((float)n.x, (float)n.y, (float)n.z, kind);
  }
}

// ----------------------------------------------------------------------------- kernels
// MLoD projection of the sorted queries q[0..m) in place (projectCuda, Fastapss.cu:398-437):
// n_iter steps per query, one lane per query; queries whose caller index perm[i] >= limit are
// skipped.  A wavefront's 64 lanes hold 64 consecutive (Morton-ordered) queries, so they start
// their walks together and load the same nodes at the same iterations for much of them.
template <int FORM, int SEM, int KT, bool CNT>
__global__ __launch_bounds__(PBS) void k_project(TreeView tv, KParams kp, float4* __restrict__ q, float4* __restrict__ qn,
                                                 const uint32_t* __restrict__ perm, uint32_t m, uint32_t limit,
                                                 uint32_t nIter, float stepMax, float cutoff, uint32_t minDepth,
                                                 unsigned long long* __restrict__ counters) {
  using T = typename FormT<FORM>::T;
  __shared__ uint32_t sFirst[kLevels][PBS];
  __shared__ uint8_t sMask[kLevels][PBS];
  __shared__ T sMul[kLevels][PBS];
  __shared__ T sDist[kLevels][PBS];
  const int tid = threadIdx.x;
  const uint32_t i = xcd_block() * PBS + tid;
  const bool live = i < m && !(limit < m && perm[i] >= limit);
  Counters ct{0, 0, 0, 0, 0};
  if (live) {
    const LevelState<T> st{sFirst, sMask, sMul, sDist};
    const float4 q4 = q[i];
    V3<double> qd{q4.x, q4.y, q4.z};
    float4 nout = make_float4(0.f, 0.f, 0.f, 0.f);
    for (uint32_t it = 0; it < nIter; ++it) {
      typename FormT<FORM>::Acc acc;
      acc.init();
      const V3<float> qf{(float)qd.x, (float)qd.y, (float)qd.z};
      V3<T> qt;
      if constexpr (FORM == 2)
        qt = V3<T>{qd.x, qd.y, qd.z};
      else
        qt = V3<T>{qf.x, qf.y, qf.z};
      traverse<FORM, SEM, KT, CNT>(acc, kp, tv, qt, (T)cutoff, minDepth, st, tid, ct);
      fit_step<FORM, SEM>(acc, qd, qf, stepMax, nout);
    }
    q[i] = make_float4((float)qd.x, (float)qd.y, (float)qd.z, 0.f);
    qn[i] = nout;
  }
  if constexpr (CNT) {
    // lane slots: 64 x the longest loop of the wavefront (lanes wait for it)
    uint32_t wmax = ct.iters;
    for (int o = 32; o > 0; o >>= 1) wmax = max(wmax, (uint32_t)__shfl_xor((int)wmax, o));
    if (live) {
      atomicAdd(&counters[0], (unsigned long long)ct.far_);
      atomicAdd(&counters[1], (unsigned long long)ct.leafPts);
      atomicAdd(&counters[2], (unsigned long long)ct.tested);
      atomicAdd(&counters[3], (unsigned long long)ct.nearN);
      atomicAdd(&counters[4], (unsigned long long)ct.iters);
    }
    if (__lane_id() == 0) atomicAdd(&counters[5], (unsigned long long)wmax);
  }
}

// k_project with the wave-shared walk: every lane of a wavefront calls the walk (lanes without a
// query take part in no node), so the walk's control flow stays wave-uniform.
template <int FORM, int SEM, int KT, bool CNT>
__global__ __launch_bounds__(PBS) void k_project_wave(TreeView tv, KParams kp, float4* __restrict__ q,
                                                      float4* __restrict__ qn, const uint32_t* __restrict__ perm,
                                                      uint32_t m, uint32_t limit, uint32_t nIter, float stepMax,
                                                      float cutoff, uint32_t minDepth,
                                                      unsigned long long* __restrict__ counters) {
  using T = typename FormT<FORM>::T;
  __shared__ WaveLevel sLv[kLevels][PBS / 64];
  __shared__ T sMul[kLevels][PBS];
  __shared__ T sDist[kLevels][PBS];
  __shared__ uint8_t sNear[kLevels][PBS];
  const int tid = threadIdx.x;
  const int wv = tid >> 6;
  const uint32_t i = xcd_block() * PBS + tid;
  const bool live = i < m && !(limit < m && perm[i] >= limit);
  Counters ct{0, 0, 0, 0, 0};
  uint32_t waveIters = 0;
  const WaveStack<T> st{sLv, sMul, sDist, sNear};
  const float4 q4 = live ? q[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  V3<double> qd{q4.x, q4.y, q4.z};
  float4 nout = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t it = 0; it < nIter; ++it) {
    typename FormT<FORM>::Acc acc;
    acc.init();
    const V3<float> qf{(float)qd.x, (float)qd.y, (float)qd.z};
    V3<T> qt;
    if constexpr (FORM == 2)
      qt = V3<T>{qd.x, qd.y, qd.z};
    else
      qt = V3<T>{qf.x, qf.y, qf.z};
    traverse_wave<FORM, SEM, KT, CNT>(acc, kp, tv, qt, (T)cutoff, minDepth, live, st, tid, wv, ct, waveIters);
    if constexpr (FORM == 1) acc.flush();
    if (live) fit_step<FORM, SEM>(acc, qd, qf, stepMax, nout);
  }
  if (live) {
    q[i] = make_float4((float)qd.x, (float)qd.y, (float)qd.z, 0.f);
    qn[i] = nout;
  }
  if constexpr (CNT) {
    if (live) {
      atomicAdd(&counters[0], (unsigned long long)ct.far_);
      atomicAdd(&counters[1], (unsigned long long)ct.leafPts);
      atomicAdd(&counters[2], (unsigned long long)ct.tested);
      atomicAdd(&counters[3], (unsigned long long)ct.nearN);
      atomicAdd(&counters[4], (unsigned long long)ct.iters);
    }
    if (__lane_id() == 0) atomicAdd(&counters[5], (unsigned long long)waveIters);
  }
}

// Global APSS over all points (GlobalAPSS.cu:113-192 / APSSStats, CPUUtils.cpp:515-528):
// point tiles staged through LDS, every lane of the workgroup sums the tile for its query.
// The fit uses the 1e-12 plane threshold of GlobalAPSS.cu:143 (the REF_CPU fit).
template <int FORM, int KT>
__global__ __launch_bounds__(PBS) void k_global(TreeView tv, KParams kp, float4* __restrict__ q, float4* __restrict__ qn,
                                                const uint32_t* __restrict__ perm, uint32_t m, uint32_t limit,
                                                uint32_t nIter, float stepMax) {
  using T = typename FormT<FORM>::T;
  __shared__ float4 sP[PBS], sN[PBS];
  const int tid = threadIdx.x;
  const uint32_t i = blockIdx.x * PBS + tid;
  const bool active = i < m && !(limit < m && perm[i] >= limit);
  const float4 q4 = active ? q[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  V3<double> qd{q4.x, q4.y, q4.z};
  float4 nout = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t it = 0; it < nIter; ++it) {
    typename FormT<FORM>::Acc acc;
    acc.init();
    const V3<float> qf{(float)qd.x, (float)qd.y, (float)qd.z};
    V3<T> qt;
    if constexpr (FORM == 2)
      qt = V3<T>{qd.x, qd.y, qd.z};
    else
      qt = V3<T>{qf.x, qf.y, qf.z};
    for (uint32_t base = 0; base < tv.n; base += PBS) {
      __syncthreads();
      if (base + tid < tv.n) {
        sP[tid] = tv.pts[base + tid];
        sN[tid] = tv.nrm[base + tid];
      }
      __syncthreads();
      const uint32_t cnt = min((uint32_t)PBS, tv.n - base);
      if (active) {
        TreeView lv = tv;
        lv.pts = sP;
        lv.nrm = sN;
        for (uint32_t j = 0; j < cnt; ++j) {
          add_point<FORM, KT>(acc, kp, lv, j, qt, T(1));
          if constexpr (FORM == 1) {
            if ((j & (kFlush - 1)) == kFlush - 1) acc.flush();
          }
        }
        if constexpr (FORM == 1) acc.flush();
      }
    }
    fit_step<FORM, MLOD_SEM_REF_CPU>(acc, qd, qf, stepMax, nout);
  }
  if (active) {
    q[i] = make_float4((float)qd.x, (float)qd.y, (float)qd.z, 0.f);
    qn[i] = nout;
  }
}

// Serving step launch: workgroup b projects the kBatchWG pool slots of pool block blocks[b].y
// with the tree, kernel and parameters of run descriptor blocks[b].x, read by the workgroup from
// the launch table with wave-uniform (scalar) loads; per-lane work and results are those of
// k_project, so a query gives the same bits alone, in any mix, and whether its steps run in one
// launch or several (the FP64 form carries its FP64 position between launches in qdp, as
// k_project carries it between the steps of one launch).  A launch for interactive requests
// raises its wavefronts' issue priority (s_setprio), so they keep close to their unloaded speed
// beside bulk launches.  A run taking its last steps writes its results to the host staging and
// flags each of its workgroups done (system-scope release).
template <int FORM, int SEM, int KT>
__global__ __launch_bounds__(PBS) void k_project_batch(const ReqDesc* __restrict__ reqs, const uint2* __restrict__ blocks,
                                                       float4* __restrict__ q, float4* __restrict__ qn,
                                                       double4* __restrict__ qdp, BatchOut bo, uint32_t prio) {
  static_assert(PBS == kBatchWG, "one workgroup per pool block");
  using T = typename FormT<FORM>::T;
  __shared__ uint32_t sFirst[kLevels][PBS];
  __shared__ uint8_t sMask[kLevels][PBS];
  __shared__ T sMul[kLevels][PBS];
  __shared__ T sDist[kLevels][PBS];
  if (prio) __builtin_amdgcn_s_setprio(3);
  const int tid = threadIdx.x;
  const uint32_t lb = xcd_block();
  const uint2 b = blocks[lb];
  const uint32_t r = __builtin_amdgcn_readfirstlane(b.x), pb = __builtin_amdgcn_readfirstlane(b.y);
  const ReqDesc& R = reqs[r];
  const uint32_t i = pb * PBS + tid, outOff = R.outOff;
  if (i < R.pad0 + R.count) {
    const TreeView tv = R.tv;
    const KParams& kp = R.kp;
    const float cutoff = R.cutoff, stepMax = R.stepMax;
    const uint32_t minDepth = R.minDepth, nIter = R.nIter;
    const LevelState<T> st{sFirst, sMask, sMul, sDist};
    V3<double> qd;
    if constexpr (FORM == 2) {
      const double4 d4 = qdp[i];
      qd = V3<double>{d4.x, d4.y, d4.z};
    } else {
      const float4 q4 = q[i];
      qd = V3<double>{q4.x, q4.y, q4.z};
    }
    float4 nout = make_float4(0.f, 0.f, 0.f, 0.f);
    Counters ct{0, 0, 0, 0, 0};
    for (uint32_t it = 0; it < nIter; ++it) {
      typename FormT<FORM>::Acc acc;
      acc.init();
      const V3<float> qf{(float)qd.x, (float)qd.y, (float)qd.z};
      V3<T> qt;
      if constexpr (FORM == 2)
        qt = V3<T>{qd.x, qd.y, qd.z};
      else
        qt = V3<T>{qf.x, qf.y, qf.z};
      traverse<FORM, SEM, KT, false>(acc, kp, tv, qt, (T)cutoff, minDepth, st, tid, ct);
      fit_step<FORM, SEM>(acc, qd, qf, stepMax, nout);
    }
    const float4 p4 = make_float4((float)qd.x, (float)qd.y, (float)qd.z, 0.f);
    q[i] = p4;
    qn[i] = nout;
    if constexpr (FORM == 2) qdp[i] = make_double4(qd.x, qd.y, qd.z, 0.0);
    if (outOff != kNoOut) {
      const uint64_t d = outOff + bo.pidx[i], e = (uint64_t)bo.cap + d;
      bo.out[3 * d] = p4.x;
      bo.out[3 * d + 1] = p4.y;
      bo.out[3 * d + 2] = p4.z;
      bo.out[3 * e] = nout.x;
      bo.out[3 * e + 1] = nout.y;
      bo.out[3 * e + 2] = nout.z;
    }
  }
  if (outOff != kNoOut) {  // uniform over the workgroup
    __threadfence_system();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(bo.flag + lb, bo.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

template <int FORM, int SEM>
void launch_batch_all(const ReqDesc* r, const uint2* blocks, uint32_t nblk, float4* q, float4* qn, double4* qd,
                      const BatchOut& bo, uint32_t prio, hipStream_t s) {
  k_project_batch<FORM, SEM, -1><<<nblk, PBS, 0, s>>>(r, blocks, q, qn, qd, bo, prio);
}

template <int FORM, int SEM, int KT, bool CNT>
int launch_one(const mlod_tree* t, const KParams& kp, uint32_t limit, uint32_t nIter, float stepMax,
               unsigned long long* counters, hipStream_t s) {
  const uint32_t m = t->qs.m;
  if (m == 0) return MLOD_OK;
  if (t->traversal == MLOD_TRAVERSAL_WAVE)
    k_project_wave<FORM, SEM, KT, CNT><<<(m + PBS - 1) / PBS, PBS, 0, s>>>(
        t->view(), kp, t->qs.q, t->qs.qn, t->qs.perm, m, limit, nIter, stepMax, t->cutoff, t->min_depth, counters);
  else
    k_project<FORM, SEM, KT, CNT><<<(m + PBS - 1) / PBS, PBS, 0, s>>>(t->view(), kp, t->qs.q, t->qs.qn, t->qs.perm, m,
                                                                       limit, nIter, stepMax, t->cutoff, t->min_depth,
                                                                       counters);
  return MLOD_OK;
}

template <int FORM, int KT>
void launch_global(const mlod_tree* t, const KParams& kp, uint32_t limit, uint32_t nIter, float stepMax, hipStream_t s) {
  const uint32_t m = t->qs.m;
  k_global<FORM, KT><<<(m + PBS - 1) / PBS, PBS, 0, s>>>(t->view(), kp, t->qs.q, t->qs.qn, t->qs.perm, m, limit, nIter,
                                                         stepMax);
}

template <int FORM, int SEM>
int dispatch_kt(const mlod_tree* t, const KParams& kp, uint32_t limit, uint32_t nIter, float stepMax, hipStream_t s) {
  switch (kp.type) {
    case MLOD_KERNEL_GAUSSIAN: return launch_one<FORM, SEM, MLOD_KERNEL_GAUSSIAN, false>(t, kp, limit, nIter, stepMax, nullptr, s);
    case MLOD_KERNEL_SINGULAR: return launch_one<FORM, SEM, MLOD_KERNEL_SINGULAR, false>(t, kp, limit, nIter, stepMax, nullptr, s);
    default: return launch_one<FORM, SEM, MLOD_KERNEL_GAUSSIAN_MULTIPLE, false>(t, kp, limit, nIter, stepMax, nullptr, s);
  }
}

template <int FORM>
void dispatch_global(const mlod_tree* t, const KParams& kp, uint32_t limit, uint32_t nIter, float stepMax, hipStream_t s) {
  switch (kp.type) {
    case MLOD_KERNEL_GAUSSIAN: launch_global<FORM, MLOD_KERNEL_GAUSSIAN>(t, kp, limit, nIter, stepMax, s); break;
    case MLOD_KERNEL_SINGULAR: launch_global<FORM, MLOD_KERNEL_SINGULAR>(t, kp, limit, nIter, stepMax, s); break;
    default: launch_global<FORM, MLOD_KERNEL_GAUSSIAN_MULTIPLE>(t, kp, limit, nIter, stepMax, s); break;
  }
}

}  // namespace

KParams make_kparams(const mlod_kernel& k) {
  KParams p{};
  p.type = k.type;
  p.ns = k.type == MLOD_KERNEL_GAUSSIAN_MULTIPLE ? k.n_sigmas : 1;
  p.c = k.constant_shift;
  p.cp = k.constant_shift > 0.0001f ? k.constant_shift : 0.0001f;
  p.e = k.exponent;
  const double er = std::nearbyint((double)k.exponent);
  p.eint = (er == (double)k.exponent && er >= 1.0 && er <= 16.0) ? (int)er : -1;
  const double log2e = 1.4426950408889634;
  if (k.type == MLOD_KERNEL_GAUSSIAN_MULTIPLE) {
    for (int s = 0; s < k.n_sigmas && s < MLOD_MAX_SIGMAS; ++s) {
      const double sg = k.sigmas[s];
      p.a[s] = (float)(-log2e / (2.0 * sg * sg));
      p.b[s] = (float)(1.0 / (sg * sg * sg));
      p.ad[s] = -1.0 / (2.0 * sg * sg);
      p.bd[s] = 1.0 / (sg * sg * sg);
    }
  } else if (k.type == MLOD_KERNEL_GAUSSIAN) {
    const double sg = k.sigma;
    p.a[0] = (float)(-log2e / (2.0 * sg * sg));
    p.ad[0] = -1.0 / (2.0 * sg * sg);
  }
  return p;
}

int launch_project_batch(const ReqDesc* d_reqs, const uint2* d_blocks, uint32_t nblk, float4* d_q, float4* d_qn,
                         double4* d_qd, const BatchOut& bo, uint32_t prio, int semantics, int precision,
                         hipStream_t s) {
  if (nblk == 0) return MLOD_OK;
  const bool cpu = semantics == MLOD_SEM_REF_CPU;
  if (precision == MLOD_PREC_FP32C) {
    if (cpu) launch_batch_all<1, MLOD_SEM_REF_CPU>(d_reqs, d_blocks, nblk, d_q, d_qn, d_qd, bo, prio, s);
    else launch_batch_all<1, MLOD_SEM_REF_GPU>(d_reqs, d_blocks, nblk, d_q, d_qn, d_qd, bo, prio, s);
  } else if (precision == MLOD_PREC_FP32) {
    if (cpu) launch_batch_all<0, MLOD_SEM_REF_CPU>(d_reqs, d_blocks, nblk, d_q, d_qn, d_qd, bo, prio, s);
    else launch_batch_all<0, MLOD_SEM_REF_GPU>(d_reqs, d_blocks, nblk, d_q, d_qn, d_qd, bo, prio, s);
  } else {
    if (!d_qd) return fail(MLOD_ERR_INVALID, "the FP64 form needs the FP64 position pool");
    if (cpu) launch_batch_all<2, MLOD_SEM_REF_CPU>(d_reqs, d_blocks, nblk, d_q, d_qn, d_qd, bo, prio, s);
    else launch_batch_all<2, MLOD_SEM_REF_GPU>(d_reqs, d_blocks, nblk, d_q, d_qn, d_qd, bo, prio, s);
  }
  LAUNCH_CHECK();
  return MLOD_OK;
}

int launch_project(mlod_tree* t, uint32_t n, uint32_t n_iter, float step_max, hipStream_t s) {
  const KParams kp = make_kparams(t->kernel);
  const uint32_t limit = n;
  if (t->mode == MLOD_MODE_GLOBAL) {
    switch (t->precision) {
      case MLOD_PREC_FP32C: dispatch_global<1>(t, kp, limit, n_iter, step_max, s); break;
      case MLOD_PREC_FP64: dispatch_global<2>(t, kp, limit, n_iter, step_max, s); break;
      default: dispatch_global<0>(t, kp, limit, n_iter, step_max, s); break;
    }
    LAUNCH_CHECK();
    return MLOD_OK;
  }
  const bool cpu = t->semantics == MLOD_SEM_REF_CPU;
  int r;
  switch (t->precision) {
    case MLOD_PREC_FP32C:
      r = cpu ? dispatch_kt<1, MLOD_SEM_REF_CPU>(t, kp, limit, n_iter, step_max, s)
              : dispatch_kt<1, MLOD_SEM_REF_GPU>(t, kp, limit, n_iter, step_max, s);
      break;
    case MLOD_PREC_FP64:
      r = cpu ? dispatch_kt<2, MLOD_SEM_REF_CPU>(t, kp, limit, n_iter, step_max, s)
              : dispatch_kt<2, MLOD_SEM_REF_GPU>(t, kp, limit, n_iter, step_max, s);
      break;
    default:
      r = cpu ? dispatch_kt<0, MLOD_SEM_REF_CPU>(t, kp, limit, n_iter, step_max, s)
              : dispatch_kt<0, MLOD_SEM_REF_GPU>(t, kp, limit, n_iter, step_max, s);
      break;
  }
  if (r != MLOD_OK) return r;
  LAUNCH_CHECK();
  return MLOD_OK;
}

// Counter pass: one FP32C step with counting on; positions are not advanced (the step
// runs on a scratch copy of the queries).
int launch_count(mlod_tree* t, uint32_t n, unsigned long long* d_out4, hipStream_t s) {
  const KParams kp = make_kparams(t->kernel);
  const bool cpu = t->semantics == MLOD_SEM_REF_CPU;
  int r;
#define MLOD_CNT_CASE(SEMV)                                                                                  \
  switch (kp.type) {                                                                                         \
    case MLOD_KERNEL_GAUSSIAN: r = launch_one<1, SEMV, MLOD_KERNEL_GAUSSIAN, true>(t, kp, n, 1, t->radius, d_out4, s); break; \
    case MLOD_KERNEL_SINGULAR: r = launch_one<1, SEMV, MLOD_KERNEL_SINGULAR, true>(t, kp, n, 1, t->radius, d_out4, s); break; \
    default: r = launch_one<1, SEMV, MLOD_KERNEL_GAUSSIAN_MULTIPLE, true>(t, kp, n, 1, t->radius, d_out4, s); break;     \
  }
  if (cpu) {
    MLOD_CNT_CASE(MLOD_SEM_REF_CPU)
  } else {
    MLOD_CNT_CASE(MLOD_SEM_REF_GPU)
  }
#undef MLOD_CNT_CASE
  if (r != MLOD_OK) return r;
  LAUNCH_CHECK();
  return MLOD_OK;
}

}  // namespace mlod

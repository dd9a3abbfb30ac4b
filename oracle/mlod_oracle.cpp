// mlod_oracle.cpp — TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// CPU restatement of the reference MLoD/APSS projection path
// (CorentinMercier/MlodSurfaces: sources/apss/CPUUtils.cpp, Kernel.cpp, Fastapss.cu,
// GlobalAPSS.cu, sources/dual-contouring/kdTreeNN/kdTreeNN.h).  It is the checker the
// parity tests compare the HIP kernels against.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it, and only as a checker; the product
// (mlodsurfaces_amd/libmlod.so) never links or calls it.
//
// Every function cites the reference file:line it restates.  Scalar type T is float
// (reference arithmetic: glm float vectors, double exp inside the weight) or double
// (the FP64 oracle: the same algorithm with every operation in double).
//
// How each part is pinned (tests/test_oracle_*.py):
//   build / cell assignment / BFS relations / areas : bit-exact against the reference's
//       own CPU code compiled into oracle/_ref (golden fixtures in tests/golden/), and
//       brute-force invariants (every point in one leaf, leaf sizes, child order).
//   node moments  : FP64 closed forms (root = sum over all points; node = sum of children).
//   weights       : closed forms against numpy.
//   MLoD traversal: golden vectors of the reference CPU path (REF_CPU semantics, FP32);
//                   cutoff -> infinity reduces it exactly to brute-force global APSS.
//   REF_GPU terms : the parent share a_c / a_node equals REF_CPU's 1 / #children on an
//                   octree of equal sibling areas (a lattice), and is checked term by term
//                   on a hand-written three-point tree; the sum over a multi-point leaf
//                   equals global APSS at cutoff -> infinity (REF_CPU falls short); the
//                   1e-6 plane threshold puts a sphere of curvature radius 6e5 on its
//                   tangent plane under REF_GPU and on the sphere under REF_CPU.
//   sphere fit    : noise-free sphere / plane data reproduce the exact projection.
//   kNN / ball    : brute force against the reference kd-tree's golden vectors.
//   adaptive octree driver : golden vectors of the reference's OctreeGenerator steps
//                   restated here (cell enclosing, neighbour expansion) plus invariants.

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <queue>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------------------
// small vector type (stands in for glm::vec3 / glm::dvec3)
// glm dot order is (x*x' + y*y') + z*z'  (lib/glm/glm/detail/func_geometric.inl:48-55)
// glm normalize is v * (1 / sqrt(dot(v, v)))  (func_geometric.inl, inversesqrt)
// ---------------------------------------------------------------------------------------
template <class T>
struct V3 {
  T x, y, z;
};
template <class T> static inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> static inline V3<T> operator+(V3<T> a, V3<T> b) { return mk<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> static inline V3<T> operator-(V3<T> a, V3<T> b) { return mk<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> static inline V3<T> operator-(V3<T> a) { return mk<T>(-a.x, -a.y, -a.z); }
template <class T> static inline V3<T> operator*(T s, V3<T> a) { return mk<T>(s * a.x, s * a.y, s * a.z); }
template <class T> static inline V3<T> operator*(V3<T> a, T s) { return mk<T>(a.x * s, a.y * s, a.z * s); }
template <class T> static inline V3<T> operator/(V3<T> a, T s) { return mk<T>(a.x / s, a.y / s, a.z / s); }
template <class T> static inline T dot(V3<T> a, V3<T> b) {
  T x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
  return (x + y) + z;
}
template <class T> static inline T length2(V3<T> a) { return dot(a, a); }
template <class T> static inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> static inline V3<T> normalize(V3<T> a) { return a * (T(1) / std::sqrt(dot(a, a))); }
template <class T> static inline bool anynan(V3<T> a) { return std::isnan(a.x) || std::isnan(a.y) || std::isnan(a.z); }

// ---------------------------------------------------------------------------------------
// Weight kernel (sources/apss/Kernel.h:37-46, Kernel.cpp:64-83)
// type numbering as Kernel.h:37: GAUSSIAN = 0, SINGULAR = 1, GAUSSIAN_MULTIPLE = 2
// ---------------------------------------------------------------------------------------
}  // namespace orc

extern "C" {
typedef struct {
  int32_t type;
  float sigma;
  float exponent;
  float constant_shift;
  int32_t n_sigmas;
  float sigmas[32];
} orc_kernel;

typedef struct {
  float cutoff;          // scalingProtectionSphere (CPUUtils.h:206)
  uint32_t min_depth;    // m_minDepth (CPUUtils.h:207)
  int32_t semantics;     // 0 = REF_GPU (Fastapss.cu), 1 = REF_CPU (CPUUtils.cpp)
  int32_t mode;          // 0 = MLoD, 1 = global APSS, 2 = kNN APSS, 3 = ball APSS
  uint32_t knn;          // k (CPUUtils.h:211)
  float ball_radius;     // r; r < 0 => 3 * sigmas[0] (CPUUtils.cpp:706-707)
} orc_params;
}

namespace orc {

// Reference-arithmetic weight (float vectors, double exp), Kernel.cpp:64-83.
static inline float weight_ref(const orc_kernel& k, V3<float> p, V3<float> eta) {
  const float d2 = length2(eta - p);  // glm::distance2(p, eta) = length2(eta - p) (gtx/norm.inl:40-44)
  if (k.type == 2) {
    double sum_g = 0.0;
    for (int i = 0; i < k.n_sigmas; ++i) {
      const float s = k.sigmas[i];
      sum_g += std::exp(-(double)d2 / (2.0 * s * s)) / (double)(s * s * s);
    }
    if (sum_g == 0) sum_g = k.constant_shift;  // Kernel.cpp:69-73 (message omitted)
    return (float)sum_g;
  }
  if (k.type == 0) return (float)(std::exp(-(double)d2 / (2.0 * k.sigma * k.sigma)) + k.constant_shift);
  if (k.type == 1) {
    const float val = std::max(k.constant_shift, 0.0001f);
    return std::pow(val, k.exponent) / std::pow(d2 + k.constant_shift, k.exponent);
  }
  return 1.0f;
}

// FP64 weight: the same formulas evaluated in double.
static inline double weight_ref(const orc_kernel& k, V3<double> p, V3<double> eta) {
  const double d2 = length2(eta - p);
  if (k.type == 2) {
    double sum_g = 0.0;
    for (int i = 0; i < k.n_sigmas; ++i) {
      const double s = k.sigmas[i];
      sum_g += std::exp(-d2 / (2.0 * s * s)) / (s * s * s);
    }
    if (sum_g == 0) sum_g = k.constant_shift;
    return sum_g;
  }
  if (k.type == 0) return std::exp(-d2 / (2.0 * (double)k.sigma * (double)k.sigma)) + (double)k.constant_shift;
  if (k.type == 1) {
    const double val = std::max((double)k.constant_shift, (double)0.0001f);
    return std::pow(val, (double)k.exponent) / std::pow(d2 + (double)k.constant_shift, (double)k.exponent);
  }
  return 1.0;
}

// Blending between levels, CPUUtils.cpp:400-408 (EXPFUNC off, Fastapss.h:27).
static inline float blend(float x) { return (float)((1 - std::cos(3.14159265 * x)) / 2.0); }
static inline double blend(double x) { return (1 - std::cos(3.14159265 * x)) / 2.0; }

// ---------------------------------------------------------------------------------------
// Moments: apssNodeStats (CPUUtils.h:29-49) and apssStats (CPUUtils.h:51-61)
// ---------------------------------------------------------------------------------------
template <class T>
struct Mom {  // per point / per node: a, a p, a n, a p.n, a |p|^2
  T a = 0;
  V3<T> ap{0, 0, 0}, an{0, 0, 0};
  T apn = 0, app = 0;
};
template <class T>
struct Stats {  // weighted sums
  V3<T> wp{0, 0, 0}, wn{0, 0, 0};
  T w = 0, wpn = 0, wpp = 0;
};
// operator*(apssNodeStats, float)  CPUUtils.cpp:86-95
template <class T> static inline Stats<T> scale(const Mom<T>& m, T w) {
  Stats<T> s;
  s.w = w * m.a;
  s.wn = w * m.an;
  s.wp = w * m.ap;
  s.wpn = w * m.apn;
  s.wpp = w * m.app;
  return s;
}
// apssStats::operator*(float)  CPUUtils.cpp:62-70
template <class T> static inline Stats<T> scale(const Stats<T>& m, T f) {
  Stats<T> s;
  s.wp = f * m.wp;
  s.wn = f * m.wn;
  s.w = m.w * f;
  s.wpn = m.wpn * f;
  s.wpp = m.wpp * f;
  return s;
}
// apssStats::operator+ / +=  CPUUtils.cpp:36-52
template <class T> static inline Stats<T> add(const Stats<T>& a, const Stats<T>& b) {
  Stats<T> s;
  s.wp = a.wp + b.wp;
  s.wn = a.wn + b.wn;
  s.w = a.w + b.w;
  s.wpn = a.wpn + b.wpn;
  s.wpp = a.wpp + b.wpp;
  return s;
}

// ---------------------------------------------------------------------------------------
// Octree (CPUUtils.h:108-161) restated with flat node storage
// ---------------------------------------------------------------------------------------
struct Node {
  int depth = 0;
  float bb[3], BB[3];
  int child[8];
  int numChildren = 0;
  std::vector<uint32_t> idx;  // leaf point indices (stable input order)
  int bfs = -1;
  Node() {
    for (int c = 0; c < 8; ++c) child[c] = -1;
  }
  bool leaf() const { return !idx.empty(); }
  V3<float> center() const {  // BBOX::center, CPUUtils.h:86
    return mk<float>((bb[0] + BB[0]) / 2.f, (bb[1] + BB[1]) / 2.f, (bb[2] + BB[2]) / 2.f);
  }
  float diag() const {  // BBOX::diagonal, CPUUtils.h:84
    V3<float> d = mk<float>(bb[0] - BB[0], bb[1] - BB[1], bb[2] - BB[2]);  // distance2(BB, bb) = length2(bb - BB)
    return std::sqrt(length2(d));
  }
  float radius() const { return diag() / 2.f; }  // CPUUtils.h:85
};

struct Tree {
  std::vector<V3<float>> P, N;
  std::vector<Node> nodes;  // creation order
  std::vector<int> bfs;     // BFS index -> node id
  std::vector<int32_t> father, firstChild, nextBrother;
  std::vector<float> area;
  std::vector<Mom<float>> ptF, ndF;   // per point, per node (node id order)
  std::vector<Mom<double>> ptD, ndD;
  float bb[3], BB[3];
  uint64_t ties = 0;
  int tie_policy = 0;
  uint32_t maxDepth = 12, maxPts = 1;
};

// octreeNode::buildRecursive, CPUUtils.cpp:271-329.  The split is (bb + BB) / 2.0 where the
// sum of the two float bounds is a float addition (CPUUtils.cpp:286-288), so the split is
// the float-rounded midpoint.  Exact ties p == split (CPUUtils.cpp:299-308) draw rand() in
// the reference; tie_policy 0 sends them to the low child, 1 to the high child, 2 replays
// glibc rand() > RAND_MAX/2 in the reference's consumption order (points in node order,
// axes x, y, z, nodes depth-first with children created in xi, yi, zi loop order).
static void buildRec(Tree& t, int id, std::vector<uint32_t> ind, unsigned depth) {
  t.nodes[id].depth = (int)depth;
  if (ind.size() <= t.maxPts || depth == t.maxDepth) {
    t.nodes[id].idx = std::move(ind);
    return;
  }
  std::vector<uint32_t> ch[8];
  double split[3], vals[3][3];
  for (int a = 0; a < 3; ++a) {
    const float s = t.nodes[id].bb[a] + t.nodes[id].BB[a];
    split[a] = (double)s / 2.0;
    vals[a][0] = t.nodes[id].bb[a];
    vals[a][1] = split[a];
    vals[a][2] = t.nodes[id].BB[a];
  }
  for (uint32_t pIdx : ind) {
    const float pc[3] = {t.P[pIdx].x, t.P[pIdx].y, t.P[pIdx].z};
    unsigned bits[3];
    for (int a = 0; a < 3; ++a) {
      bits[a] = 0;
      if ((double)pc[a] == split[a]) {
        ++t.ties;
        if (t.tie_policy == 1) bits[a] = 1;
        if (t.tie_policy == 2 && rand() > RAND_MAX / 2) bits[a] = 1;
      }
      if ((double)pc[a] > split[a]) bits[a] = 1;
    }
    ch[bits[0] + 2 * bits[1] + 4 * bits[2]].push_back(pIdx);
  }
  int nc = 0;
  for (unsigned xi = 0; xi < 2; ++xi)
    for (unsigned yi = 0; yi < 2; ++yi)
      for (unsigned zi = 0; zi < 2; ++zi) {
        const unsigned c = xi + 2 * yi + 4 * zi;
        if (ch[c].empty()) continue;
        ++nc;
        const int cid = (int)t.nodes.size();
        t.nodes.emplace_back();
        t.nodes[id].child[c] = cid;
        Node& cn = t.nodes[cid];
        const unsigned b[3] = {xi, yi, zi};
        for (int a = 0; a < 3; ++a) {  // bbox.set(vec3(xVals[xi],..), vec3(xVals[xi+1],..)) CPUUtils.cpp:321-322
          cn.bb[a] = (float)vals[a][b[a]];
          cn.BB[a] = (float)vals[a][b[a] + 1];
        }
        buildRec(t, cid, std::move(ch[c]), depth + 1);
      }
  t.nodes[id].numChildren = nc;
}

// buildapssNodeStatsRecursive, CPUUtils.cpp:358-398
template <class T>
static void momentsRec(Tree& t, int id, const std::vector<Mom<T>>& pts, std::vector<Mom<T>>& out) {
  Mom<T> m;
  const Node& nd = t.nodes[id];
  auto acc = [&m](const Mom<T>& o) {
    m.a += o.a;
    m.an = m.an + o.an;
    m.ap = m.ap + o.ap;
    m.apn += o.apn;
    m.app += o.app;
  };
  if (nd.leaf()) {
    for (uint32_t i : nd.idx) acc(pts[i]);
  } else {
    for (int c = 0; c < 8; ++c)
      if (nd.child[c] >= 0) {
        momentsRec(t, nd.child[c], pts, out);
        acc(out[nd.child[c]]);
      }
  }
  out[id] = m;
}

// APSSOctree::build, CPUUtils.cpp:540-569 (computeBoundingBox 571-579, estimateAreas 581-587,
// buildapssNodeStats 589-599, index_nodes_by_breadth_first_search 601-619, getRelations 213-234)
// box6: a fixed root box {bb, BB} (build with updateRootBoundingBox = false, CPUUtils.cpp:547), or
// NULL for the points' bounding box.
static Tree* build(const float* xyz, const float* nrm, uint32_t n, uint32_t maxDepth, uint32_t maxPts, int tie_policy,
                   const float* box6) {
  Tree* t = new Tree();
  t->maxDepth = maxDepth;
  t->maxPts = maxPts;
  t->tie_policy = tie_policy;
  t->P.resize(n);
  t->N.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    t->P[i] = mk<float>(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    t->N[i] = mk<float>(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  }
  if (box6) {
    for (int a = 0; a < 3; ++a) t->bb[a] = box6[a], t->BB[a] = box6[3 + a];
  } else {
    for (int a = 0; a < 3; ++a) t->bb[a] = t->BB[a] = (&t->P[0].x)[a];
    for (uint32_t i = 1; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        const float v = (&t->P[i].x)[a];
        t->bb[a] = std::min(t->bb[a], v);
        t->BB[a] = std::max(t->BB[a], v);
      }
  }
  t->nodes.reserve(2 * (size_t)n + 16);
  t->nodes.emplace_back();
  for (int a = 0; a < 3; ++a) {
    t->nodes[0].bb[a] = t->bb[a];
    t->nodes[0].BB[a] = t->BB[a];
  }
  std::vector<uint32_t> ind(n);
  for (uint32_t i = 0; i < n; ++i) ind[i] = i;
  buildRec(*t, 0, std::move(ind), 0);

  // estimateAreasRecursive, CPUUtils.cpp:236-252
  t->area.assign(n, 0.f);
  for (const Node& nd : t->nodes)
    if (nd.leaf()) {
      const float d = nd.diag();
      const float leafArea = d * d;
      const float sampleArea = leafArea / (float)nd.idx.size();
      for (uint32_t i : nd.idx) t->area[i] = sampleArea;
    }
  // per point moments: apssNodeStats(p, n, a), CPUUtils.h:45
  t->ptF.resize(n);
  t->ptD.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    const float a = t->area[i];
    const V3<float> p = t->P[i], q = t->N[i];
    Mom<float>& f = t->ptF[i];
    f.ap = a * p;
    f.an = a * q;
    f.apn = a * dot(p, q);
    f.app = a * dot(p, p);
    f.a = a;
    const double ad = a;
    const V3<double> pd = mk<double>(p.x, p.y, p.z), qd = mk<double>(q.x, q.y, q.z);
    Mom<double>& d = t->ptD[i];
    d.ap = ad * pd;
    d.an = ad * qd;
    d.apn = ad * dot(pd, qd);
    d.app = ad * dot(pd, pd);
    d.a = ad;
  }
  t->ndF.resize(t->nodes.size());
  t->ndD.resize(t->nodes.size());
  momentsRec<float>(*t, 0, t->ptF, t->ndF);
  momentsRec<double>(*t, 0, t->ptD, t->ndD);

  // BFS indexing (CPUUtils.cpp:601-619) and relations (CPUUtils.cpp:213-234)
  std::queue<int> qu;
  qu.push(0);
  while (!qu.empty()) {
    const int id = qu.front();
    qu.pop();
    t->nodes[id].bfs = (int)t->bfs.size();
    t->bfs.push_back(id);
    for (int c = 0; c < 8; ++c)
      if (t->nodes[id].child[c] >= 0) qu.push(t->nodes[id].child[c]);
  }
  const size_t nn = t->nodes.size();
  t->father.assign(nn, -1);
  t->firstChild.assign(nn, -1);
  t->nextBrother.assign(nn, -1);
  for (size_t b = 0; b < nn; ++b) {
    const Node& nd = t->nodes[t->bfs[b]];
    int prev = -1;
    for (int c = 0; c < 8; ++c)
      if (nd.child[c] >= 0) {
        const int cb = t->nodes[nd.child[c]].bfs;
        t->father[cb] = (int)b;
        if (prev < 0)
          t->firstChild[b] = cb;
        else
          t->nextBrother[prev] = cb;
        prev = cb;
      }
  }
  return t;
}

// ---------------------------------------------------------------------------------------
// Sphere fit + projection: APSSOctree::projectOnAPSSStats, CPUUtils.cpp:739-782 (REF_CPU,
// plane switch 1e-12); Fastapss.cu:354-395 (REF_GPU, plane switch 1e-6; normalize(0) = 0).
// ---------------------------------------------------------------------------------------
// `aliased` restates the kNN / ball call projectOnAPSSStats(outputPoint, outputPoint, ...)
// (CPUUtils.cpp:735): q and outputPoint are one object there, so the NaN guard's
// `outputPoint = q` keeps the NaN position (normal still zeroed) and the step length
// |outputPoint - q| is always 0, i.e. no step clamp.
template <class T>
static void projectOnStats(V3<T> q, V3<T>& outP, V3<T>& outN, const Stats<T>& s, T stepMax, bool refGpu,
                           bool aliased = false) {
  const T A = s.wpn - dot(s.wp, s.wn / s.w);
  const T B = s.wpp - dot(s.wp, s.wp / s.w);
  const T u4 = (T)((0.5 * (double)A) / (double)B);
  const V3<T> u123 = (s.wn - (T(2) * u4) * s.wp) / s.w;
  const T u0 = -(dot(s.wp, u123) + u4 * s.wpp) / s.w;
  V3<T> p = q, nrm;
  const double tau = refGpu ? 0.000001 : 0.000000000001;
  auto nz = [refGpu](V3<T> v) {
    if (refGpu && dot(v, v) == T(0)) return mk<T>(0, 0, 0);
    return normalize(v);
  };
  if (std::fabs((double)u4) < tau) {
    const V3<T> n = -u123;
    const T lambda = (u0 - dot(p, n)) / length2(n);
    p = p + lambda * n;
    nrm = nz(s.wn);
  } else {
    const V3<T> c = u123 / (T(-2) * u4);
    const T r = std::sqrt(std::max<T>(T(0), length2(c) - u0 / u4));
    const V3<T> pc = nz(p - c);
    p = c + r * pc;
    nrm = nz(u123 + (T(2) * u4) * p);
  }
  if (anynan(p)) {
    outP = aliased ? p : q;
    outN = mk<T>(0, 0, 0);
    return;
  }
  if (aliased) {
    outP = p;
    outN = nrm;
    return;
  }
  const T step = length(p - q);
  if (step > stepMax) p = q + (stepMax / step) * (p - q);
  outP = p;
  outN = nrm;
}

// ---------------------------------------------------------------------------------------
// MLoD traversal.  REF_CPU: approximateAPSSStatsRecursive, CPUUtils.cpp:443-513 (a leaf keeps
// only its last point, 455-456; parent share 1/#children, 485).  REF_GPU:
// approximateAPSSNonRecursive, Fastapss.cu:160-333 (a leaf sums its points, 189-200; parent
// share a_child/a_node, 244).
// ---------------------------------------------------------------------------------------
template <class T> static inline V3<T> cvt(V3<float> v) { return mk<T>(v.x, v.y, v.z); }
template <class T> static inline const std::vector<Mom<T>>& ptMom(const Tree& t);
template <> inline const std::vector<Mom<float>>& ptMom<float>(const Tree& t) { return t.ptF; }
template <> inline const std::vector<Mom<double>>& ptMom<double>(const Tree& t) { return t.ptD; }
template <class T> static inline const std::vector<Mom<T>>& ndMom(const Tree& t);
template <> inline const std::vector<Mom<float>>& ndMom<float>(const Tree& t) { return t.ndF; }
template <> inline const std::vector<Mom<double>>& ndMom<double>(const Tree& t) { return t.ndD; }

template <class T>
static Stats<T> approx(const Tree& t, int id, V3<T> q, const orc_kernel& k, unsigned minDepth, T cutoff, bool refCpu) {
  Stats<T> acc;
  const Node& nd = t.nodes[id];
  const std::vector<Mom<T>>& pm = ptMom<T>(t);
  const std::vector<Mom<T>>& nm = ndMom<T>(t);
  if (nd.leaf()) {
    for (uint32_t i : nd.idx) {
      const T w = weight_ref(k, pm[i].ap / pm[i].a, q);
      if (refCpu)
        acc = scale(pm[i], w);
      else
        acc = add(acc, scale(pm[i], w));
    }
    return acc;
  }
  if ((unsigned)nd.depth >= minDepth) {
    const Mom<T>& m = nm[id];
    const V3<T> p = m.ap / m.a;
    const V3<T> c = cvt<T>(nd.center());
    const T r = (T)nd.radius();
    const T dq = length(q - c);
    if (dq > cutoff * r) return scale(m, weight_ref(k, p, q));
    const T wq = weight_ref(k, p, q);
    const T distToParent = std::fabs(dq - (r * cutoff));
    for (int ci = 0; ci < 8; ++ci) {
      const int cid = nd.child[ci];
      if (cid < 0) continue;
      const Node& cn = t.nodes[cid];
      const T share = refCpu ? wq / T(nd.numChildren) : wq * nm[cid].a / m.a;
      const Stats<T> parentStats = scale(m, share);
      const T dc = std::max<T>(length(q - cvt<T>(cn.center())) - ((T)cn.radius() * cutoff), T(0));
      const T wChild = distToParent / (distToParent + dc);
      const Stats<T> childStats = approx<T>(t, cid, q, k, minDepth, cutoff, refCpu);
      const T valChild = blend(wChild);
      const T valParent = 1 - valChild;
      acc = add(acc, add(scale(childStats, valChild), scale(parentStats, valParent)));
    }
    return acc;
  }
  for (int ci = 0; ci < 8; ++ci)
    if (nd.child[ci] >= 0) acc = add(acc, approx<T>(t, nd.child[ci], q, k, minDepth, cutoff, refCpu));
  return acc;
}

// octreeNode::APSSStats, CPUUtils.cpp:515-528 (global APSS over all points, area weights)
template <class T>
static Stats<T> globalStats(const Tree& t, V3<T> q, const orc_kernel& k) {
  Stats<T> s;
  const std::vector<Mom<T>>& pm = ptMom<T>(t);
  for (size_t i = 0; i < pm.size(); ++i) {
    const T w = weight_ref(k, pm[i].ap / pm[i].a, q);
    s.w = w * pm[i].a + s.w;
    s.wn = s.wn + w * pm[i].an;
    s.wp = s.wp + w * pm[i].ap;
    s.wpn = w * pm[i].apn + s.wpn;
    s.wpp = w * pm[i].app + s.wpp;
  }
  return s;
}

// Neighbour search with the kd-tree's distance arithmetic (kdTreeNN.h:416-461, 463-493):
// float component differences and squares, summed in float, widened to double.
static inline double kdSqrDist(V3<float> p, V3<float> q) {
  const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
  const float s = dx * dx + dy * dy;
  return (double)(s + dz * dz);
}
// exact k nearest (kdTree::knearest, kdTreeNN.h:590-602): k smallest squared distances;
// equal distances ordered by point index.
static void knnBrute(const Tree& t, V3<float> q, unsigned k, std::vector<uint32_t>& out) {
  const size_t n = t.P.size();
  std::vector<std::pair<double, uint32_t>> d(n);
  for (size_t i = 0; i < n; ++i) d[i] = {kdSqrDist(t.P[i], q), (uint32_t)i};
  const size_t kk = std::min<size_t>(k, n);
  std::partial_sort(d.begin(), d.begin() + kk, d.end());
  out.resize(kk);
  for (size_t i = 0; i < kk; ++i) out[i] = d[i].second;
}
// inclusive ball (kdTree::getInBall, kdTreeNN.h:581-586 / 463-493): d^2 <= r^2 in double.
static void ballBrute(const Tree& t, V3<float> q, double r, std::vector<uint32_t>& out) {
  out.clear();
  for (size_t i = 0; i < t.P.size(); ++i)
    if (kdSqrDist(t.P[i], q) <= r * r) out.push_back((uint32_t)i);
}

// One query through n_iter steps.  MLoD / global: APSSOctree::projectCPU, CPUUtils.cpp:621-636
// with oneStepProjection 784-793.  kNN / ball: projectAPSSCPU, CPUUtils.cpp:690-737.
template <class T>
static void projectOne(const Tree& t, const orc_kernel& k, const orc_params& prm, V3<float> q0, unsigned nIter, float stepMax,
                       V3<T>& outP, V3<T>& outN) {
  const bool refCpu = prm.semantics == 1;
  const T sm = stepMax < 0 ? (T)[&] {
    Node root = t.nodes[0];
    return root.radius();
  }()
                           : (T)stepMax;
  V3<T> q = cvt<T>(q0);
  outP = q;
  outN = mk<T>(0, 0, 0);
  if (prm.mode == 2 || prm.mode == 3) {
    std::vector<uint32_t> nb;
    for (unsigned it = 0; it < nIter; ++it) {
      const V3<float> qf = mk<float>((float)outP.x, (float)outP.y, (float)outP.z);
      if (prm.mode == 2) {
        knnBrute(t, qf, prm.knn, nb);
      } else {
        const double r = prm.ball_radius < 0.f ? (double)(k.sigmas[0] * 3) : (double)prm.ball_radius;
        ballBrute(t, qf, r, nb);
        if (nb.empty()) {  // CPUUtils.cpp:711-716
          outP = cvt<T>(q0);
          outN = mk<T>(0, 0, 0);
          return;
        }
      }
      Stats<T> s;
      for (uint32_t i : nb) {  // CPUUtils.cpp:721-731: unit area, raw p and n
        const V3<T> p = cvt<T>(t.P[i]), n = cvt<T>(t.N[i]);
        const T wi = weight_ref(k, p, outP);
        s.wpn += wi * dot(p, n);
        s.wpp += wi * dot(p, p);
        s.wp = s.wp + wi * p;
        s.wn = s.wn + wi * n;
        s.w += wi;
      }
      projectOnStats<T>(outP, outP, outN, s, sm, false, true);
    }
    return;
  }
  for (unsigned it = 0; it < nIter; ++it) {
    Stats<T> s = prm.mode == 1 ? globalStats<T>(t, q, k) : approx<T>(t, 0, q, k, prm.min_depth, (T)prm.cutoff, refCpu);
    projectOnStats<T>(q, outP, outN, s, sm, prm.mode == 0 && !refCpu);
    q = outP;
  }
}

}  // namespace orc

// ---------------------------------------------------------------------------------------
// C API (ctypes)
// ---------------------------------------------------------------------------------------
extern "C" {

void* orc_tree_build(const float* xyz, const float* nrm, uint32_t n, uint32_t max_depth, uint32_t max_pts_per_leaf,
                     int32_t tie_policy, uint32_t seed, const float* box6) {
  if (n == 0) return nullptr;
  if (tie_policy == 2) srand(seed);  // the reference's rand() stream (ref_harness seeds it the same way)
  return orc::build(xyz, nrm, n, max_depth, max_pts_per_leaf, tie_policy, box6);
}
void orc_tree_free(void* h) { delete static_cast<orc::Tree*>(h); }
uint32_t orc_tree_num_nodes(void* h) { return (uint32_t) static_cast<orc::Tree*>(h)->nodes.size(); }
uint64_t orc_tree_ties(void* h) { return static_cast<orc::Tree*>(h)->ties; }
void orc_tree_bbox(void* h, float* out6) {
  auto* t = static_cast<orc::Tree*>(h);
  for (int a = 0; a < 3; ++a) {
    out6[a] = t->bb[a];
    out6[3 + a] = t->BB[a];
  }
}

// BFS-ordered node table.  leaf_offset indexes orc_tree_leaf_indices' output.
void orc_tree_nodes(void* h, int32_t* depth, float* bb6, float* center3, float* radius, int32_t* father, int32_t* first_child,
                    int32_t* next_brother, int32_t* num_children, uint32_t* leaf_count, uint64_t* leaf_offset) {
  auto* t = static_cast<orc::Tree*>(h);
  uint64_t off = 0;
  for (size_t b = 0; b < t->bfs.size(); ++b) {
    const orc::Node& nd = t->nodes[t->bfs[b]];
    depth[b] = nd.depth;
    for (int a = 0; a < 3; ++a) {
      bb6[6 * b + a] = nd.bb[a];
      bb6[6 * b + 3 + a] = nd.BB[a];
    }
    const orc::V3<float> c = nd.center();
    center3[3 * b] = c.x;
    center3[3 * b + 1] = c.y;
    center3[3 * b + 2] = c.z;
    radius[b] = nd.radius();
    father[b] = t->father[b];
    first_child[b] = t->firstChild[b];
    next_brother[b] = t->nextBrother[b];
    num_children[b] = nd.numChildren;
    leaf_count[b] = (uint32_t)nd.idx.size();
    leaf_offset[b] = off;
    off += nd.idx.size();
  }
}
void orc_tree_leaf_indices(void* h, uint32_t* out) {
  auto* t = static_cast<orc::Tree*>(h);
  size_t o = 0;
  for (int id : t->bfs)
    for (uint32_t i : t->nodes[id].idx) out[o++] = i;
}
void orc_tree_areas(void* h, float* out) {
  auto* t = static_cast<orc::Tree*>(h);
  std::memcpy(out, t->area.data(), t->area.size() * sizeof(float));
}
// Node moments in BFS order, 9 values: a, ap(3), an(3), apn, app.
void orc_tree_node_moments(void* h, float* outF, double* outD) {
  auto* t = static_cast<orc::Tree*>(h);
  for (size_t b = 0; b < t->bfs.size(); ++b) {
    const auto& f = t->ndF[t->bfs[b]];
    const auto& d = t->ndD[t->bfs[b]];
    if (outF) {
      const float v[9] = {f.a, f.ap.x, f.ap.y, f.ap.z, f.an.x, f.an.y, f.an.z, f.apn, f.app};
      std::memcpy(outF + 9 * b, v, sizeof(v));
    }
    if (outD) {
      const double v[9] = {d.a, d.ap.x, d.ap.y, d.ap.z, d.an.x, d.an.y, d.an.z, d.apn, d.app};
      std::memcpy(outD + 9 * b, v, sizeof(v));
    }
  }
}

// Weight of one pair (for the closed-form weight tests).
double orc_weight(const orc_kernel* k, const float* p, const float* q, int32_t precision) {
  if (precision == 1)
    return orc::weight_ref(*k, orc::mk<double>(p[0], p[1], p[2]), orc::mk<double>(q[0], q[1], q[2]));
  return orc::weight_ref(*k, orc::mk<float>(p[0], p[1], p[2]), orc::mk<float>(q[0], q[1], q[2]));
}

// Projects m queries through n_iter steps.  precision 0 = reference arithmetic (float),
// 1 = FP64.  Outputs are widened to double.  step_max < 0 => root bbox radius.
int32_t orc_project(void* h, const orc_kernel* k, const orc_params* prm, int32_t precision, const float* q, uint32_t m,
                    uint32_t n_iter, float step_max, double* out_p, double* out_n, int32_t nthreads) {
  auto* t = static_cast<orc::Tree*>(h);
  if (!t || !k || !prm) return -1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < (int64_t)m; ++i) {
    const orc::V3<float> qi = orc::mk<float>(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
    if (precision == 1) {
      orc::V3<double> p, n;
      orc::projectOne<double>(*t, *k, *prm, qi, n_iter, step_max, p, n);
      out_p[3 * i] = p.x, out_p[3 * i + 1] = p.y, out_p[3 * i + 2] = p.z;
      out_n[3 * i] = n.x, out_n[3 * i + 1] = n.y, out_n[3 * i + 2] = n.z;
    } else {
      orc::V3<float> p, n;
      orc::projectOne<float>(*t, *k, *prm, qi, n_iter, step_max, p, n);
      out_p[3 * i] = p.x, out_p[3 * i + 1] = p.y, out_p[3 * i + 2] = p.z;
      out_n[3 * i] = n.x, out_n[3 * i + 1] = n.y, out_n[3 * i + 2] = n.z;
    }
  }
  return 0;
}

// The weighted sums of one MLoD (mode 0) or global (mode 1) pass for m queries, in FP64:
// out9 per query = {sum w, sum w p (3), sum w n (3), sum w p.n, sum w |p|^2} (apssStats,
// CPUUtils.h:51-61), before the fit.  Used to check the traversal sums term by term.
int32_t orc_stats(void* h, const orc_kernel* k, const orc_params* prm, const float* q, uint32_t m, double* out9) {
  auto* t = static_cast<orc::Tree*>(h);
  if (!t || !k || !prm || (prm->mode != 0 && prm->mode != 1)) return -1;
  for (uint32_t i = 0; i < m; ++i) {
    const orc::V3<double> qi = orc::mk<double>(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
    const orc::Stats<double> s = prm->mode == 1 ? orc::globalStats<double>(*t, qi, *k)
                                                : orc::approx<double>(*t, 0, qi, *k, prm->min_depth, (double)prm->cutoff,
                                                                      prm->semantics == 1);
    const double v[9] = {s.w, s.wp.x, s.wp.y, s.wp.z, s.wn.x, s.wn.y, s.wn.z, s.wpn, s.wpp};
    std::memcpy(out9 + 9 * (size_t)i, v, sizeof(v));
  }
  return 0;
}

// Fitted algebraic sphere for one set of weighted sums (closed-form tests of the fit).
void orc_fit(const double* sums9, const double* q, double step_max, int32_t ref_gpu, double* out_p, double* out_n) {
  orc::Stats<double> s;
  s.w = sums9[0];
  s.wp = orc::mk<double>(sums9[1], sums9[2], sums9[3]);
  s.wn = orc::mk<double>(sums9[4], sums9[5], sums9[6]);
  s.wpn = sums9[7];
  s.wpp = sums9[8];
  orc::V3<double> p, n;
  orc::projectOnStats<double>(orc::mk<double>(q[0], q[1], q[2]), p, n, s, step_max, ref_gpu != 0);
  out_p[0] = p.x, out_p[1] = p.y, out_p[2] = p.z;
  out_n[0] = n.x, out_n[1] = n.y, out_n[2] = n.z;
}

}  // extern "C"

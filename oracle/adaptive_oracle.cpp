// adaptive_oracle.cpp — TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// CPU restatement of the reference's adaptive-octree driver, OctreeGenerator::generate
// (CorentinMercier/MlodSurfaces sources/scene/upsample.cpp:228-251) and
// OctreeGenerator::generateFromPoints (upsample.cpp:128-144), over the octree-id helpers of
// sources/dual-contouring/octree.h.  The projection step (find_stable_nodesGPU,
// upsample.cpp:317-346) calls a caller-supplied callback, so the tests can drive the same
// bookkeeping with the oracle's projection, with an analytic projection, or with the GPU
// projection (checking the GPU driver's set logic bit for bit).
//
// The reference's octree.h / upsample.cpp need Eigen, which this image lacks, so they are not
// compiled here; this restatement is pinned by an independent set-based restatement with an
// analytic sphere projection and by closed forms (parent = floor division, centre / enclosing
// round trip), in tests/test_oracle_adaptive.py.
//
// Node order everywhere is the reference's node_comparator (octree.h:147-154): depth, then z,
// y, x, signed.

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <vector>

namespace orc_adaptive {

// octree_id (octree.h:18-44): depth and signed integer position
struct Id {
  int d, x, y, z;
  bool operator==(const Id& o) const { return d == o.d && x == o.x && y == o.y && z == o.z; }
};

// node_comparator, octree.h:147-154
static bool less(const Id& a, const Id& b) {
  if (a.d != b.d) return a.d < b.d;
  if (a.z != b.z) return a.z < b.z;
  if (a.y != b.y) return a.y < b.y;
  return a.x < b.x;
}

// remove_duplicates, upsample.cpp:28-32
static void dedup(std::vector<Id>& v) {
  std::sort(v.begin(), v.end(), less);
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

// one halving step of octree_id::parent / ancestor_at_depth (octree.h:547-573): truncating
// division, then one less for odd negative coordinates
static int half(int p) {
  const int t = p / 2;
  return (p < 0 && p - 2 * t != 0) ? t - 1 : t;
}

static Id ancestor(Id id, int d) {  // ancestor_at_depth, octree.h:561-573
  while (id.d > d) {
    id.d--;
    id.x = half(id.x), id.y = half(id.y), id.z = half(id.z);
  }
  return id;
}

static Id child(const Id& id, unsigned c) {  // octree.h:575-583
  return Id{id.d + 1, 2 * id.x + int(c & 1), 2 * id.y + int((c >> 1) & 1), 2 * id.z + int((c >> 2) & 1)};
}

static Id neighbor(const Id& id, unsigned n) {  // octree.h:591-602
  const int z = n / 9, y = (n - 9 * z) / 3, x = n - 9 * z - 3 * y;
  return Id{id.d, id.x + x - 1, id.y + y - 1, id.z + z - 1};
}

// expand, upsample.cpp:409-419: the 26 neighbours (octree_id_neighbors skips index 13, the
// node itself) of every node at `depth` (all nodes when depth < 0)
static void expand(std::vector<Id>& v, int depth) {
  const size_t n = v.size();
  for (size_t t = 0; t < n; ++t) {
    if (depth >= 0 && v[t].d != depth) continue;
    for (unsigned k = 0; k < 27; ++k)
      if (k != 13) v.push_back(neighbor(v[t], k));
  }
  dedup(v);
}

// aabb_octree over [-1, 1]^3 (upsample.cpp:319-322; octree.h:939-965) and over a given box
struct Box {
  float mn[3], mx[3];
};

static void center(const Box& b, const Id& id, float out[3]) {  // aabb_octree::center, octree.h:952-955
  for (int a = 0; a < 3; ++a) {
    const float diag = b.mx[a] - b.mn[a];
    const float w = diag / (float)(1 << id.d);
    const int p = a == 0 ? id.x : a == 1 ? id.y : id.z;
    const float lo = b.mn[a] + w * (float)p;
    const float hi = b.mn[a] + w * (float)(p + 1);
    out[a] = (lo + hi) / 2.0f;
  }
}

// aabb_octree::enclosing, octree.h:957-965.  A non-finite coordinate (the reference then
// converts NaN / inf to int, undefined in C++) is reported through *bad.
static Id enclosing(const Box& b, const float pos[3], int depth, bool* bad) {
  const float w = 1.0f / (float)(1 << depth);
  int c[3];
  for (int a = 0; a < 3; ++a) {
    const float diag = b.mx[a] - b.mn[a];
    const float P = ((pos[a] - b.mn[a]) / diag) / w;
    const float f = std::floor(P);
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) {
      *bad = true;
      c[a] = 0;
    } else {
      c[a] = (int)f;
    }
  }
  return Id{depth, c[0], c[1], c[2]};
}

typedef void (*proj_cb)(const float* xyz, uint32_t n, float* out_pts, float* out_nmls, void* user);

struct Gen {
  Box box{{-1.f, -1.f, -1.f}, {1.f, 1.f, 1.f}};
  proj_cb cb = nullptr;
  void* user = nullptr;
  std::vector<Id> leaves, total;
  uint32_t maxRel = 0;
  uint64_t projected = 0, calls = 0, bad = 0;
};

// find_stable_nodesGPU, upsample.cpp:317-346: project every node centre once with an
// unbounded step; keep the cell of the projection at the node's depth, or the root when the
// projected normal is degenerate (|n|^2 < 0.5)
static void find_stable(Gen& g, const std::vector<Id>& nodes, std::vector<Id>& out) {
  const size_t n = nodes.size();
  std::vector<float> c(3 * n), p(3 * n), nm(3 * n);
  for (size_t i = 0; i < n; ++i) center(g.box, nodes[i], &c[3 * i]);
  g.cb(c.data(), (uint32_t)n, p.data(), nm.data(), g.user);
  g.projected += n;
  g.calls += 1;
  out.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const float* q = &nm[3 * i];
    const float l2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];  // glm::length2
    if (l2 < 0.5f) {
      out[i] = Id{0, 0, 0, 0};
    } else {
      bool bad = false;
      out[i] = enclosing(g.box, &p[3 * i], nodes[i].d, &bad);
      if (bad) {
        out[i] = Id{0, 0, 0, 0};
        g.bad += 1;
      }
    }
  }
  dedup(out);
}

// compute_leaves, upsample.cpp:274-287
static std::vector<Id> compute_leaves(Gen& g, std::vector<Id> cur, unsigned maxDepth, std::vector<Id>* total) {
  std::vector<Id> sub;
  for (unsigned depth = 0; depth < maxDepth + 1; ++depth) {
    find_stable(g, cur, sub);
    cur.clear();
    cur.reserve(sub.size() * 8);
    for (const Id& s : sub)  // add_children, upsample.cpp:348-358
      for (unsigned c = 0; c < 8; ++c) cur.push_back(child(s, c));
    sub.clear();
    expand(cur, (int)depth + 1);
    if (total) total->insert(total->end(), cur.begin(), cur.end());
  }
  return cur;
}

// getRelativeDepth, upsample.cpp:146-161 (min / max persist across calls)
static int relative_depth(const std::vector<Id>& nodes, int mx[3], int mn[3]) {
  if (nodes.empty()) return 0;
  for (const Id& n : nodes) {
    const Id a = ancestor(n, 0);
    const int p[3] = {a.x, a.y, a.z};
    for (int k = 0; k < 3; ++k) {
      mn[k] = std::min(mn[k], p[k]);
      mx[k] = std::max(mx[k], p[k]);
    }
  }
  int dd = 0;
  for (int k = 0; k < 3; ++k) {
    const int ext = mx[k] - mn[k] + 1;
    dd = std::max(dd, (int)std::ceil(float(std::log2(ext))));
  }
  return dd;
}

// onlySetDepthNodes, upsample.cpp:163-173 (with cut, upsample.cpp:95-100)
static void only_set_depth(std::vector<Id>& nodes, unsigned depth) {
  for (Id& n : nodes)
    if ((unsigned)n.d > depth) n = ancestor(n, (int)depth);
  nodes.erase(std::remove_if(nodes.begin(), nodes.end(), [&](const Id& n) { return n.d > (int)depth; }), nodes.end());
  dedup(nodes);
}

static std::vector<Id> depth0_ancestors(std::vector<Id> v) {  // upsample.cpp:360-368
  for (Id& n : v) n = ancestor(n, 0);
  dedup(v);
  return v;
}

// OctreeGenerator::generate, upsample.cpp:228-251
static void generate(Gen& g, unsigned maxDepth) {
  std::vector<Id> depth0{Id{0, 0, 0, 0}}, old0, nodes;
  std::vector<Id>& total = g.total;
  total.clear();
  unsigned maxRel = maxDepth;
  int mn[3] = {INT_MAX, INT_MAX, INT_MAX};
  int mx[3] = {-INT_MAX, -INT_MAX, -INT_MAX};
  while (!depth0.empty()) {
    total.insert(total.end(), depth0.begin(), depth0.end());
    std::vector<Id> cur = compute_leaves(g, depth0, maxRel, &total);
    maxRel = (unsigned)std::max(0, (int)maxDepth - relative_depth(cur, mx, mn));
    nodes.insert(nodes.end(), cur.begin(), cur.end());
    old0.insert(old0.end(), depth0.begin(), depth0.end());
    depth0 = depth0_ancestors(cur);
    expand(depth0, -1);
    // vector_difference_inplace, upsample.cpp:102-111
    depth0.erase(std::remove_if(depth0.begin(), depth0.end(),
                                [&](const Id& n) { return std::find(old0.begin(), old0.end(), n) != old0.end(); }),
                 depth0.end());
  }
  dedup(nodes);
  only_set_depth(nodes, maxRel);
  only_set_depth(total, maxRel);
  g.leaves = nodes;
  g.maxRel = maxRel;
}

}  // namespace orc_adaptive

using namespace orc_adaptive;

extern "C" {

void* orc_gen_create(proj_cb cb, void* user) {
  Gen* g = new Gen();
  g->cb = cb;
  g->user = user;
  return g;
}
void orc_gen_free(void* h) { delete static_cast<Gen*>(h); }

void orc_gen_run(void* h, uint32_t max_depth) { generate(*static_cast<Gen*>(h), max_depth); }

// generateFromPoints, upsample.cpp:128-144: the cells at max_depth enclosing the points, in
// the given box (NULL: [-1, 1]^3)
void orc_gen_from_points(void* h, const float* xyz, uint32_t n, uint32_t max_depth, const float* box6) {
  Gen& g = *static_cast<Gen*>(h);
  Box b = g.box;
  if (box6) b = Box{{box6[0], box6[1], box6[2]}, {box6[3], box6[4], box6[5]}};
  g.leaves.assign(n, Id{0, 0, 0, 0});
  g.total.clear();
  for (uint32_t i = 0; i < n; ++i) {
    bool bad = false;
    g.leaves[i] = enclosing(b, xyz + 3 * (size_t)i, (int)max_depth, &bad);
    if (bad) g.bad += 1;
  }
  dedup(g.leaves);
  g.maxRel = max_depth;
}

// counts: [leaves, total, max relative depth, centres projected, projection calls, bad]
void orc_gen_info(void* h, uint64_t* out6) {
  const Gen& g = *static_cast<Gen*>(h);
  out6[0] = g.leaves.size();
  out6[1] = g.total.size();
  out6[2] = g.maxRel;
  out6[3] = g.projected;
  out6[4] = g.calls;
  out6[5] = g.bad;
}

// ids as (depth, x, y, z) rows; which = 0 leaves, 1 total
void orc_gen_ids(void* h, int32_t which, int32_t* out) {
  const Gen& g = *static_cast<Gen*>(h);
  const std::vector<Id>& v = which ? g.total : g.leaves;
  for (size_t i = 0; i < v.size(); ++i) {
    out[4 * i] = v[i].d;
    out[4 * i + 1] = v[i].x;
    out[4 * i + 2] = v[i].y;
    out[4 * i + 3] = v[i].z;
  }
}

// aabb_octree::center / enclosing of single ids (for the closed-form tests)
void orc_cell_center(const int32_t* id4, const float* box6, float* out3) {
  Box b{{-1.f, -1.f, -1.f}, {1.f, 1.f, 1.f}};
  if (box6) b = Box{{box6[0], box6[1], box6[2]}, {box6[3], box6[4], box6[5]}};
  center(b, Id{id4[0], id4[1], id4[2], id4[3]}, out3);
}
void orc_cell_ancestor(const int32_t* id4, int32_t depth, int32_t* out4) {
  const Id a = ancestor(Id{id4[0], id4[1], id4[2], id4[3]}, depth);
  out4[0] = a.d, out4[1] = a.x, out4[2] = a.y, out4[3] = a.z;
}

}  // extern "C"

// dc_oracle.cpp — TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// CPU restatement of the reference's dual-contouring consumer of the projection, for checking
// the GPU implementation (mlodsurfaces_amd/csrc/dc.hip):
//   * uniformize (sources/scene/upsample.cpp:262-272): depth0_ancestors (360-367),
//     retreiveCorrectRoot / union_nodes (369-406), grade (67-93), regularize_up (40-46), cut
//     (95-100), regularize_siblings (48-65), then sorted_octree (octree.h:232-312); and
//     uniformizeFromPoints (253-259): grade, regularize_up, regularize_siblings only;
//   * dual_contour::extract (sources/dual-contouring/dual_contouring.h:59-87) over
//     minimal_edges::traverse (octree.h:400-537), with compute_vertex (130-176) and the corner
//     values of scene::computeDualContouring (scene.cpp:1346-1396) given by the caller.
// The reference's octree.h / dual_contouring.h / upsample.cpp need Eigen, which this image lacks, so
// they cannot be built here; this restatement follows them step by step in plain C++ (the
// operations on floats in the same order; QEF with the zero normal field the reference passes,
// i.e. the mean of the edge intersections).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

struct Id {
  int d, x, y, z;
  bool operator==(const Id& o) const { return d == o.d && x == o.x && y == o.y && z == o.z; }
};

// node_comparator (octree.h:147-154)
bool less_id(const Id& a, const Id& b) {
  if (a.d != b.d) return a.d < b.d;
  if (a.z != b.z) return a.z < b.z;
  if (a.y != b.y) return a.y < b.y;
  return a.x < b.x;
}

// octree_id::parent (octree.h:546-558): division toward zero, minus one for odd negatives
int half(int p) { return p / 2 - ((p < 0 && (p - 2 * (p / 2)) != 0) ? 1 : 0); }
Id parent(const Id& n) { return Id{n.d - 1, half(n.x), half(n.y), half(n.z)}; }
Id ancestor0(Id n) {
  while (n.d > 0) n = parent(n);
  return n;
}
Id child(const Id& n, unsigned c) { return Id{n.d + 1, 2 * n.x + int(c & 1), 2 * n.y + int((c >> 1) & 1), 2 * n.z + int((c >> 2) & 1)}; }

void remove_duplicates(std::vector<Id>& v) {
  std::sort(v.begin(), v.end(), less_id);
  v.erase(std::unique(v.begin(), v.end()), v.end());
}
void remove_duplicates_reverse(std::vector<Id>& v) {
  std::sort(v.rbegin(), v.rend(), less_id);
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

// grade (upsample.cpp:67-93)
void grade(std::vector<Id>& nodes) {
  static const unsigned corners[8] = {0, 2, 6, 8, 18, 20, 24, 26};
  std::sort(nodes.rbegin(), nodes.rend(), less_id);
  size_t i = 0;
  int current = nodes.empty() ? 0 : nodes[0].d;
  while (i < nodes.size()) {
    if (current != nodes[i].d) {
      remove_duplicates_reverse(nodes);
      current = nodes[i].d;
    }
    if (nodes[i].d > 0) {
      for (unsigned c : corners) {
        const int z = c / 9, y = (c - 9 * z) / 3, x = c - 9 * z - 3 * y;  // octree_id::neighbor
        const Id n = nodes[i];
        nodes.push_back(parent(Id{n.d, n.x + x - 1, n.y + y - 1, n.z + z - 1}));
      }
    }
    i++;
  }
  remove_duplicates(nodes);
}

struct Tree {
  std::vector<Id> nodes;           // sorted by node_comparator
  std::vector<int32_t> kids;        // 8 per node: the children's indices (children are not contiguous), -1 for a leaf
  int maxDepth = 0;
  float diag = 1.f, origin[3] = {0.f, 0.f, 0.f};  // sorted octree -> unit frame (retreiveCorrectRoot)
  // corner values and the extracted mesh
  std::unordered_map<uint64_t, float> corners;
  std::vector<float> verts;       // 3 per vertex, sorted-octree unit frame
  std::vector<uint32_t> vertNode;  // node of each vertex
  std::vector<uint32_t> tris;

  long find(const Id& n) const {
    auto it = std::lower_bound(nodes.begin(), nodes.end(), n, less_id);
    return (it != nodes.end() && *it == n) ? long(it - nodes.begin()) : -1;
  }
};

uint64_t corner_key(long x, long y, long z) {
  return ((uint64_t)(uint32_t)z << 42) | ((uint64_t)(uint32_t)y << 21) | (uint64_t)(uint32_t)x;
}

void uniformize(Tree& T, std::vector<Id> nodes, unsigned maxDepthCut, bool doCut, bool fromPoints = false) {
  if (fromPoints) {  // uniformizeFromPoints: no re-rooting, no cut
    T.diag = 1.f;
    doCut = false;
  } else {
  // depth0_ancestors, appended
  std::vector<Id> d0(nodes.size());
  for (size_t i = 0; i < nodes.size(); ++i) d0[i] = ancestor0(nodes[i]);
  remove_duplicates(d0);
  nodes.insert(nodes.end(), d0.begin(), d0.end());
  // retreiveCorrectRoot / union_nodes
  int mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {-INT32_MAX, -INT32_MAX, -INT32_MAX};
  bool has0 = false;
  for (const Id& n : nodes) {
    if (n.d != 0) continue;
    const int p[3] = {n.x, n.y, n.z};
    for (int a = 0; a < 3; ++a) mn[a] = std::min(mn[a], p[a]), mx[a] = std::max(mx[a], p[a]);
    has0 = true;
  }
  int delta = 0;
  if (has0)
    for (int a = 0; a < 3; ++a) delta = std::max(delta, (int)std::ceil(float(std::log2(mx[a] - mn[a] + 1))));
  if (has0 && delta > 0) {
    for (Id& n : nodes) {
      n.x -= mn[0] * (1 << n.d), n.y -= mn[1] * (1 << n.d), n.z -= mn[2] * (1 << n.d);
      n.d += delta;
    }
    T.diag = float(1 << delta);
    for (int a = 0; a < 3; ++a) T.origin[a] = float(mn[a]);
  } else {
    // union_nodes returns depth -1: no shift; the reference's 1 << -1 is undefined, taken as 1
    T.diag = 1.f;
  }
  }
  grade(nodes);
  // regularize_up (upsample.cpp:40-46)
  for (size_t t = 0; t < nodes.size(); ++t)
    if (nodes[t].d != 0) nodes.push_back(parent(nodes[t]));
  remove_duplicates(nodes);
  if (doCut)
    nodes.erase(std::remove_if(nodes.begin(), nodes.end(), [&](const Id& n) { return n.d > int(maxDepthCut); }),
                nodes.end());
  // regularize_siblings (upsample.cpp:48-65)
  const size_t s = nodes.size();
  nodes.resize(9 * s);
  for (size_t t = 0; t < s; ++t) {
    const Id n = nodes[t];
    if (n.d <= 0) {  // the slots keep a default-constructed id: the root
      for (unsigned k = 0; k < 8; ++k) nodes[s + 8 * t + k] = Id{0, 0, 0, 0};
      continue;
    }
    const Id p = parent(n);
    for (unsigned k = 0; k < 8; ++k) nodes[s + 8 * t + k] = child(p, k);
  }
  remove_duplicates(nodes);
  T.nodes = nodes;
  T.maxDepth = 0;
  for (const Id& n : nodes) T.maxDepth = std::max(T.maxDepth, n.d);
  T.kids.assign(8 * nodes.size(), -1);
  for (size_t t = 0; t < nodes.size(); ++t)
    for (unsigned k = 0; k < 8; ++k) T.kids[8 * t + k] = int32_t(T.find(child(nodes[t], k)));
}

enum Dir { X = 0, Y = 1, Z = 2 };

struct Extract {
  Tree& T;
  std::vector<int64_t> nodeVertex;

  explicit Extract(Tree& t) : T(t), nodeVertex(t.nodes.size(), -1) {}

  bool leaf(size_t n) const { return T.kids[8 * n] < 0; }
  uint32_t child_of(size_t n, unsigned c) const { return uint32_t(T.kids[8 * n + c]); }

  float get_corner(long x, long y, long z) const {
    auto it = T.corners.find(corner_key(x, y, z));
    return it == T.corners.end() ? std::nanf("") : it->second;
  }

  // compute_vertex (dual_contouring.h:130-176) with the zero normal field
  void compute_vertex(size_t ni, float out[3]) const {
    static const uint8_t edges[12][2] = {{0, 1}, {0, 2}, {1, 3}, {2, 3}, {4, 5}, {4, 6},
                                         {5, 7}, {6, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    const Id& n = T.nodes[ni];
    const float w = 1.0f / float(1 << T.maxDepth);
    const int delta = 1 << (T.maxDepth - n.d);
    float sum[3] = {0.f, 0.f, 0.f};
    size_t count = 0;
    for (const auto& e : edges) {
      const long u[3] = {long(delta) * (n.x + (e[0] & 1)), long(delta) * (n.y + ((e[0] >> 1) & 1)),
                         long(delta) * (n.z + ((e[0] >> 2) & 1))};
      const long v[3] = {long(delta) * (n.x + (e[1] & 1)), long(delta) * (n.y + ((e[1] >> 1) & 1)),
                         long(delta) * (n.z + ((e[1] >> 2) & 1))};
      const float fu = get_corner(u[0], u[1], u[2]), fv = get_corner(v[0], v[1], v[2]);
      if ((fu < 0) == (fv < 0)) continue;
      const float x = -fu / (fv - fu);
      for (int a = 0; a < 3; ++a) {
        const volatile float xv = x * (float(v[a]) * w);
        const volatile float xu = (1 - x) * (float(u[a]) * w);
        sum[a] += xv + xu;
      }
      ++count;
    }
    const float width = 1.0f / float(1 << n.d);
    const float lo[3] = {float(n.x) * width, float(n.y) * width, float(n.z) * width};
    const float hi[3] = {(float(n.x) + 1.f) * width, (float(n.y) + 1.f) * width, (float(n.z) + 1.f) * width};
    float p[3];
    for (int a = 0; a < 3; ++a) p[a] = sum[a] / float(count);  // QEF centre; the zero system adds nothing
    bool inside = true, nan = false;
    for (int a = 0; a < 3; ++a) {
      inside = inside && lo[a] <= p[a] && p[a] <= hi[a];
      nan = nan || std::isnan(p[a]);
    }
    if (count == 0 || (!inside && nan) || nan) {
      const int q[3] = {n.x, n.y, n.z};
      for (int a = 0; a < 3; ++a) p[a] = (float(q[a]) + 0.5f) * width;
    }
    for (int a = 0; a < 3; ++a) out[a] = p[a] + 0.0f;
  }

  uint32_t get_vertex(size_t n) {
    if (nodeVertex[n] < 0) {
      nodeVertex[n] = int64_t(T.vertNode.size());
      float p[3];
      compute_vertex(n, p);
      T.verts.insert(T.verts.end(), p, p + 3);
      T.vertNode.push_back(uint32_t(n));
    }
    return uint32_t(nodeVertex[n]);
  }

  // on_min_edges callback (dual_contouring.h:195-233) and the triangles of extract (59-87)
  void min_edge(size_t n1, size_t n2, size_t n3, size_t n4, Dir dir) {
    const size_t nodes[4] = {n1, n2, n3, n4};
    unsigned sel = 0;
    int selDepth = -1;
    for (unsigned i = 0; i < 4; ++i)
      if (T.nodes[nodes[i]].d > selDepth) sel = i, selDepth = T.nodes[nodes[i]].d;
    static const unsigned ci[3][4][2] = {{{6, 7}, {4, 5}, {2, 3}, {0, 1}},
                                         {{7, 5}, {6, 4}, {3, 1}, {2, 0}},
                                         {{3, 7}, {2, 6}, {1, 5}, {0, 4}}};
    const Id& n = T.nodes[nodes[sel]];
    const long delta = 1L << (T.maxDepth - n.d);
    const unsigned c0 = ci[dir][sel][0], c1 = ci[dir][sel][1];
    const float f0 = get_corner(delta * (n.x + (c0 & 1)), delta * (n.y + ((c0 >> 1) & 1)), delta * (n.z + ((c0 >> 2) & 1)));
    const float f1 = get_corner(delta * (n.x + (c1 & 1)), delta * (n.y + ((c1 >> 1) & 1)), delta * (n.z + ((c1 >> 2) & 1)));
    if (std::isnan(f0) || std::isnan(f1)) return;
    const bool m0 = f0 < 0, m1 = f1 < 0;
    if (m0 == m1) return;
    // the reference's first test compares the corner numbers (never equal here); vertices are
    // created for all three nodes before the equality test on them
    auto tri = [&](unsigned i0, unsigned i1, unsigned i2) {
      const uint32_t v[3] = {get_vertex(nodes[i0]), get_vertex(nodes[i1]), get_vertex(nodes[i2])};
      if (v[0] == v[1] || v[1] == v[2] || v[2] == v[0]) return;
      T.tris.push_back(v[0]);
      T.tris.push_back(m0 ? v[2] : v[1]);
      T.tris.push_back(m0 ? v[1] : v[2]);
    };
    tri(0, 1, 3);
    tri(0, 3, 2);
  }

  // minimal_edges (octree.h:400-537)
  void node(size_t n) {
    if (leaf(n)) return;
    uint32_t c[8];
    for (unsigned k = 0; k < 8; ++k) c[k] = child_of(n, k);
    for (unsigned k = 0; k < 8; ++k) node(c[k]);
    face(c[0], c[1], X), face(c[2], c[3], X), face(c[4], c[5], X), face(c[6], c[7], X);
    face(c[0], c[2], Y), face(c[1], c[3], Y), face(c[4], c[6], Y), face(c[5], c[7], Y);
    face(c[0], c[4], Z), face(c[1], c[5], Z), face(c[2], c[6], Z), face(c[3], c[7], Z);
    edge(c[0], c[2], c[4], c[6], X), edge(c[1], c[3], c[5], c[7], X);
    edge(c[0], c[1], c[4], c[5], Y), edge(c[2], c[3], c[6], c[7], Y);
    edge(c[0], c[1], c[2], c[3], Z), edge(c[4], c[5], c[6], c[7], Z);
  }

  void face(size_t n1, size_t n2, Dir dir) {
    const bool l1 = leaf(n1), l2 = leaf(n2);
    if (l1 && l2) return;
    static const unsigned subfaces[3][2][4] = {{{1, 3, 5, 7}, {0, 2, 4, 6}},
                                               {{2, 3, 6, 7}, {0, 1, 4, 5}},
                                               {{4, 5, 6, 7}, {0, 1, 2, 3}}};
    for (unsigned f = 0; f < 4; ++f)
      face(l1 ? n1 : child_of(n1, subfaces[dir][0][f]), l2 ? n2 : child_of(n2, subfaces[dir][1][f]), dir);
    static const unsigned subedges[3][4][4] = {{{1, 5, 0, 4}, {3, 7, 2, 6}, {1, 3, 0, 2}, {5, 7, 4, 6}},
                                               {{3, 7, 1, 5}, {2, 6, 0, 4}, {2, 3, 0, 1}, {6, 7, 4, 5}},
                                               {{4, 5, 0, 1}, {6, 7, 2, 3}, {4, 6, 0, 2}, {5, 7, 1, 3}}};
    static const Dir subedgedir[3][4] = {{Y, Y, Z, Z}, {X, X, Z, Z}, {Y, Y, X, X}};
    for (unsigned e = 0; e < 4; ++e) {
      const size_t m1 = l1 ? n1 : child_of(n1, subedges[dir][e][0]);
      const size_t m2 = l1 ? n1 : child_of(n1, subedges[dir][e][1]);
      const size_t m3 = l2 ? n2 : child_of(n2, subedges[dir][e][2]);
      const size_t m4 = l2 ? n2 : child_of(n2, subedges[dir][e][3]);
      const bool flip = dir == X || (dir == Y && e < 2);
      if (!flip)
        edge(m1, m2, m3, m4, subedgedir[dir][e]);
      else
        edge(m1, m3, m2, m4, subedgedir[dir][e]);
    }
  }

  void edge(size_t n1, size_t n2, size_t n3, size_t n4, Dir dir) {
    const bool l1 = leaf(n1), l2 = leaf(n2), l3 = leaf(n3), l4 = leaf(n4);
    if (l1 && l2 && l3 && l4) {
      min_edge(n1, n2, n3, n4, dir);
      return;
    }
    static const unsigned subedges[3][2][4] = {{{6, 4, 2, 0}, {7, 5, 3, 1}},
                                               {{5, 4, 1, 0}, {7, 6, 3, 2}},
                                               {{3, 2, 1, 0}, {7, 6, 5, 4}}};
    for (unsigned e = 0; e < 2; ++e)
      edge(l1 ? n1 : child_of(n1, subedges[dir][e][0]), l2 ? n2 : child_of(n2, subedges[dir][e][1]),
           l3 ? n3 : child_of(n3, subedges[dir][e][2]), l4 ? n4 : child_of(n4, subedges[dir][e][3]), dir);
  }
};

}  // namespace

extern "C" {

// Uniformized sorted octree of the given cells (ids4: depth, x, y, z per cell).
void* orc_dc_build(const int32_t* ids4, uint32_t n, uint32_t max_depth, int32_t cut) {
  std::vector<Id> nodes(n);
  for (uint32_t i = 0; i < n; ++i) nodes[i] = Id{ids4[4 * i], ids4[4 * i + 1], ids4[4 * i + 2], ids4[4 * i + 3]};
  Tree* T = new Tree();
  uniformize(*T, nodes, max_depth, cut != 0);
  return T;
}
// uniformizeFromPoints (upsample.cpp:253-259)
void* orc_dc_build_points(const int32_t* ids4, uint32_t n) {
  std::vector<Id> nodes(n);
  for (uint32_t i = 0; i < n; ++i) nodes[i] = Id{ids4[4 * i], ids4[4 * i + 1], ids4[4 * i + 2], ids4[4 * i + 3]};
  Tree* T = new Tree();
  uniformize(*T, nodes, 0, false, true);
  return T;
}
void orc_dc_free(void* h) { delete static_cast<Tree*>(h); }
uint32_t orc_dc_num_nodes(void* h) { return uint32_t(static_cast<Tree*>(h)->nodes.size()); }
// diag, origin[3], max depth
void orc_dc_frame(void* h, float* out4, int32_t* max_depth) {
  Tree* T = static_cast<Tree*>(h);
  out4[0] = T->diag;
  for (int a = 0; a < 3; ++a) out4[1 + a] = T->origin[a];
  *max_depth = T->maxDepth;
}
void orc_dc_nodes(void* h, int32_t* ids4, uint8_t* is_leaf) {
  Tree* T = static_cast<Tree*>(h);
  for (size_t i = 0; i < T->nodes.size(); ++i) {
    const Id& n = T->nodes[i];
    ids4[4 * i] = n.d, ids4[4 * i + 1] = n.x, ids4[4 * i + 2] = n.y, ids4[4 * i + 3] = n.z;
    is_leaf[i] = T->kids[8 * i] < 0;
  }
}
// Corner values at max-depth integer corner positions (xyz3); later entries of a key win.
void orc_dc_set_corners(void* h, const int32_t* xyz3, const float* f, uint32_t m) {
  Tree* T = static_cast<Tree*>(h);
  T->corners.clear();
  T->corners.reserve(m);
  for (uint32_t i = 0; i < m; ++i) T->corners[corner_key(xyz3[3 * i], xyz3[3 * i + 1], xyz3[3 * i + 2])] = f[i];
}
// Runs the extraction; returns vertex and triangle counts.
void orc_dc_extract(void* h, uint32_t* n_verts, uint32_t* n_tris) {
  Tree* T = static_cast<Tree*>(h);
  T->verts.clear(), T->vertNode.clear(), T->tris.clear();
  Extract E(*T);
  const long root = T->find(Id{0, 0, 0, 0});
  if (root >= 0) E.node(size_t(root));
  *n_verts = uint32_t(T->vertNode.size());
  *n_tris = uint32_t(T->tris.size() / 3);
}
void orc_dc_mesh(void* h, float* verts, uint32_t* vert_node, uint32_t* tris) {
  Tree* T = static_cast<Tree*>(h);
  std::memcpy(verts, T->verts.data(), T->verts.size() * sizeof(float));
  std::memcpy(vert_node, T->vertNode.data(), T->vertNode.size() * sizeof(uint32_t));
  std::memcpy(tris, T->tris.data(), T->tris.size() * sizeof(uint32_t));
}

// mesh::smoothMesh (sources/scene/upsample.h:55-76), sequentially as the reference
void orc_mesh_smooth(float* P, uint32_t nv, const uint32_t* T, uint32_t nt) {
  std::vector<float> np(3 * (size_t)nv, 0.f), val(nv, 0.f);
  for (uint32_t t = 0; t < nt; ++t) {
    const uint32_t a = T[3 * t], b = T[3 * t + 1], c = T[3 * t + 2];
    for (int q = 0; q < 3; ++q) {
      const float ab = P[3 * b + q] + P[3 * c + q], ba = P[3 * a + q] + P[3 * c + q], ca = P[3 * b + q] + P[3 * a + q];
      np[3 * a + q] += ab;
      np[3 * b + q] += ba;
      np[3 * c + q] += ca;
    }
    val[a] += 2, val[b] += 2, val[c] += 2;
  }
  for (uint32_t v = 0; v < nv; ++v)
    for (int q = 0; q < 3; ++q) P[3 * v + q] = np[3 * v + q] / val[v];
}

// mesh::subdivideMesh (upsample.h:78-133): P holds room for nv + 3 nt vertices, T for 4 nt
// triangles; returns the new vertex count
uint32_t orc_mesh_subdivide(float* P, uint32_t nv, uint32_t* T, uint32_t nt) {
  std::unordered_map<uint64_t, uint32_t> onEdge;
  uint32_t n = nv;
  size_t tn = 3 * (size_t)nt;
  auto mid = [&](uint32_t u, uint32_t v) {
    const uint64_t k = ((uint64_t)std::min(u, v) << 32) | std::max(u, v);
    auto it = onEdge.find(k);
    if (it != onEdge.end()) return it->second;
    for (int q = 0; q < 3; ++q) P[3 * (size_t)n + q] = (P[3 * u + q] + P[3 * v + q]) / 2.f;
    onEdge[k] = n;
    return n++;
  };
  for (uint32_t t = 0; t < nt; ++t) {
    const uint32_t a = T[3 * t], b = T[3 * t + 1], c = T[3 * t + 2];
    const uint32_t ab = mid(a, b), bc = mid(c, b), ca = mid(c, a);
    T[3 * t] = a, T[3 * t + 1] = ab, T[3 * t + 2] = ca;
    const uint32_t add[9] = {ab, b, bc, ab, bc, ca, ca, bc, c};
    for (int q = 0; q < 9; ++q) T[tn++] = add[q];
  }
  return n;
}

}  // extern "C"

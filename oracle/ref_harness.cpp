// ref_harness.cpp — TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// A thin C API around the reference's own CPU path (sources/apss/CPUUtils.cpp and
// Kernel.cpp of CorentinMercier/MlodSurfaces, compiled where they lie by oracle/Makefile
// into oracle/_ref/libmlodref.so).  It reaches the protected
// APSSOctree::projectCPU(q, out, n, iters, fast, kernel, step) (CPUUtils.h:198-199) through
// a subclass, exactly as Fastapss::projectCPU / GlobalAPSS::projectCPU do
// (Fastapss.cu:513-516, GlobalAPSS.cu:283-286), and dumps the built octree in BFS order
// the way Fastapss's copy() flattens it (Fastapss.cu:97-150, via getRelations
// CPUUtils.cpp:213-234).  Used to write golden vectors (tests/golden/) and as the
// reported CPU baseline; never by the product.

#include <cstdint>
#include <cstdlib>
#include <vector>

#include "apss/CPUUtils.h"

namespace {
struct RefTree : public APSSOctree {
  using APSSOctree::projectCPU;
};
}  // namespace

extern "C" {

typedef struct {
  int32_t type;
  float sigma;
  float exponent;
  float constant_shift;
  int32_t n_sigmas;
  float sigmas[32];
} ref_kernel;

typedef struct {
  float cutoff;
  uint32_t min_depth;
  int32_t semantics;  // unused: the reference CPU path has one semantics
  int32_t mode;       // 0 = MLoD (fast), 1 = global (brute force), 2 = kNN, 3 = ball
  uint32_t knn;
  float ball_radius;
} ref_params;

// box6: a fixed root box {bb, BB}: it is set on the root node and the tree is built with
// updateRootBoundingBox = false (CPUUtils.cpp:540-547); NULL builds with the points' box.
void* ref_tree_build(const float* xyz, const float* nrm, uint32_t n, uint32_t max_depth, uint32_t max_pts, uint32_t seed,
                     const float* box6) {
  if (n == 0) return nullptr;
  srand(seed);  // exact-midpoint ties draw rand() (CPUUtils.cpp:299-308)
  std::vector<glm::vec3> P(n), N(n);
  for (uint32_t i = 0; i < n; ++i) {
    P[i] = glm::vec3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    N[i] = glm::vec3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  }
  RefTree* t = new RefTree();
  if (box6) {
    // the reference offers no setter for the root box; its root node is created non-const
    BBOX b;
    b.set(glm::vec3(box6[0], box6[1], box6[2]), glm::vec3(box6[3], box6[4], box6[5]));
    const_cast<octreeNode*>(t->get_root())->setBoundingBox(b);
  }
  t->build(P, N, box6 == nullptr, max_depth, max_pts);
  return t;
}

void ref_tree_free(void* h) { delete static_cast<RefTree*>(h); }

uint32_t ref_tree_num_nodes(void* h) { return static_cast<RefTree*>(h)->get_root()->nbOfNodes(); }

void ref_tree_nodes(void* h, int32_t* depth, float* bb6, float* center3, float* radius, int32_t* father, int32_t* first_child,
                    int32_t* next_brother, int32_t* num_children, uint32_t* leaf_count, uint64_t* leaf_offset, uint32_t* leaf_idx,
                    float* moments9) {
  RefTree* t = static_cast<RefTree*>(h);
  const uint32_t nn = t->get_root()->nbOfNodes();
  std::vector<const octreeNode*> nodes(nn, nullptr);
  std::vector<relations> rel(nn);
  t->get_root()->getRelations(nodes, rel);
  uint64_t off = 0;
  for (uint32_t b = 0; b < nn; ++b) {
    const octreeNode* nd = nodes[b];
    const BBOX& bx = nd->getBoundingBox();
    depth[b] = (int32_t)nd->depth;
    for (int a = 0; a < 3; ++a) {
      bb6[6 * b + a] = bx.bb[a];
      bb6[6 * b + 3 + a] = bx.BB[a];
    }
    const glm::vec3 c = bx.center();
    center3[3 * b] = c.x;
    center3[3 * b + 1] = c.y;
    center3[3 * b + 2] = c.z;
    radius[b] = bx.radius();
    father[b] = rel[b].father;
    first_child[b] = rel[b].firstChild;
    next_brother[b] = rel[b].nextBrother;
    num_children[b] = (int32_t)nd->numberOfChildren;
    leaf_count[b] = (uint32_t)nd->indices.size();
    leaf_offset[b] = off;
    for (unsigned i : nd->indices) leaf_idx[off++] = i;
    const apssNodeStats& s = nd->nodeapssNodeStats;
    const float v[9] = {s.s_ai, s.s_ai_pi.x, s.s_ai_pi.y, s.s_ai_pi.z, s.s_ai_ni.x, s.s_ai_ni.y, s.s_ai_ni.z, s.s_ai_pi_ni, s.s_ai_pi_pi};
    for (int j = 0; j < 9; ++j) moments9[9 * b + j] = v[j];
  }
}

void ref_tree_areas(void* h, float* out) {
  const std::vector<apssNodeStats>& s = static_cast<RefTree*>(h)->getPointsApssNodeStats();
  for (size_t i = 0; i < s.size(); ++i) out[i] = s[i].s_ai;
}

void ref_tree_bbox(void* h, float* out6) {
  const BBOX& b = static_cast<RefTree*>(h)->getBoundingBox();
  for (int a = 0; a < 3; ++a) {
    out6[a] = b.bb[a];
    out6[3 + a] = b.BB[a];
  }
}

static Kernel make_kernel(const ref_kernel* k) {
  Kernel K;
  K.type = (Kernel::KernelType)k->type;
  K.sigma = k->sigma;
  K.exponent = k->exponent;
  K.constant_shift = k->constant_shift;
  K.sigmas.assign(k->sigmas, k->sigmas + k->n_sigmas);
  return K;
}

// scene::projectPointsOnCPU, scene.cpp:968-1018: OpenMP parallel-for over queries, each one
// APSS::projectCPU call.  Returns 0.
int32_t ref_project(void* h, const ref_kernel* k, const ref_params* prm, const float* q, uint32_t m, uint32_t n_iter, float step_max,
                    float* out_p, float* out_n, int32_t nthreads) {
  RefTree* t = static_cast<RefTree*>(h);
  const Kernel K = make_kernel(k);
  t->setScalingProtectionSphere(prm->cutoff);
  t->setMinDepth(prm->min_depth);
  t->knnMode(prm->mode == 2, prm->knn);
  t->ballMode(prm->mode == 3, prm->ball_radius);
  if (prm->mode != 2) t->knnMode(false, prm->knn);
  const bool fast = prm->mode != 1;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < (int64_t)m; ++i) {
    glm::vec3 p(q[3 * i], q[3 * i + 1], q[3 * i + 2]), o, n(0.f);
    t->projectCPU(p, o, n, n_iter, fast, K, step_max);
    out_p[3 * i] = o.x, out_p[3 * i + 1] = o.y, out_p[3 * i + 2] = o.z;
    out_n[3 * i] = n.x, out_n[3 * i + 1] = n.y, out_n[3 * i + 2] = n.z;
  }
  return 0;
}

// Exact k nearest through the reference kd-tree (APSSOctree::getKnn, CPUUtils.cpp:665-672 ->
// kdTree::knearest, kdTreeNN.h:590-602), in the heap order APSSOctree sums them.
// out holds m*k*3 floats (neighbour positions).
void ref_knn(void* h, const float* q, uint32_t m, uint32_t k, float* out) {
  RefTree* t = static_cast<RefTree*>(h);
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)m; ++i) {
    std::vector<glm::vec3> nb;
    t->getKnn(glm::vec3(q[3 * i], q[3 * i + 1], q[3 * i + 2]), k, nb);
    for (uint32_t j = 0; j < k; ++j) {
      out[(3 * (size_t)i * k) + 3 * j] = nb[j].x;
      out[(3 * (size_t)i * k) + 3 * j + 1] = nb[j].y;
      out[(3 * (size_t)i * k) + 3 * j + 2] = nb[j].z;
    }
  }
}

}  // extern "C"

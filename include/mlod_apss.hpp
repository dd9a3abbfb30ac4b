// mlod_apss.hpp — header-only C++ drop-in for the reference's projection plugin interface.
//
// Same class names, method names, argument meaning and defaults as the reference's
// `APSS` (sources/apss/GPUUtils.h:22-67), `Fastapss` (Fastapss.h:32-51), `GlobalAPSS`
// (GlobalAPSS.h:19-43), `Kernel` (Kernel.h:24-60) and the parts of `APSSOctree`
// (CPUUtils.h:166-229) the projection path uses, implemented on libmlod.so (mlod.h).
// Vector arguments take any 3-float type laid out as x, y, z (glm::vec3 included).
//
// Differences: errors throw mlod::Error instead of exit() (gpuErrchk, GPUUtils.cu:32-41);
// the octree is built on the GPU by APSSOctree::build; projectCPU computes the reference's CPU
// function on the GPU (the product has no CPU path); set_semantics / set_precision choose the
// reference function and the arithmetic form (see mlod.h).
#ifndef MLOD_APSS_HPP
#define MLOD_APSS_HPP

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "mlod.h"

namespace mlod {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int code) {
  if (code != MLOD_OK) throw Error(code, std::string("libmlod: ") + mlod_last_error());
}

template <class V>
inline const float* fptr(const V* v) {
  static_assert(sizeof(V) == 3 * sizeof(float), "vector type must be three packed floats");
  return reinterpret_cast<const float*>(v);
}
template <class V>
inline float* fptr(V* v) {
  static_assert(sizeof(V) == 3 * sizeof(float), "vector type must be three packed floats");
  return reinterpret_cast<float*>(v);
}

// Kernel (Kernel.h:24-60); defaults of Kernel::Kernel (Kernel.cpp:13-16).
class Kernel {
 public:
  enum KernelType { GAUSSIAN, SINGULAR, GAUSSIAN_MULTIPLE };
  float sigma = 2.0f;
  float exponent = 1.0f;
  float constant_shift = 0.0f;
  std::vector<float> sigmas;
  std::vector<float> fixedSigmas;
  float beginningSigma = 0.0f;
  float factor = 0.0f;
  unsigned nbOfComputedSigmas = 1;
  KernelType type = GAUSSIAN;

  // Kernel::updateSigmas (Kernel.cpp:110-128).
  void updateSigmas(bool computedSigmas = true, float radius = 1.f) {
    if (computedSigmas) {
      sigmas.resize(nbOfComputedSigmas);
      sigmas[0] = beginningSigma;
      for (unsigned i = 1; i < nbOfComputedSigmas; i++) sigmas[i] = sigmas[i - 1] * factor;
    } else {
      sigmas = fixedSigmas;
    }
    sigma *= radius;
    for (float& s : sigmas) s *= radius;
  }

  mlod_kernel to_c() const {
    mlod_kernel k{};
    k.type = (int32_t)type;
    k.sigma = sigma;
    k.exponent = exponent;
    k.constant_shift = constant_shift;
    if (sigmas.size() > MLOD_MAX_SIGMAS) throw Error(MLOD_ERR_INVALID, "at most 32 sigmas");
    k.n_sigmas = (int32_t)sigmas.size();
    for (size_t i = 0; i < sigmas.size(); ++i) k.sigmas[i] = sigmas[i];
    return k;
  }
};

// The octree (APSSOctree, CPUUtils.h:166-229), built and resident on the GPU.
class APSSOctree {
 public:
  APSSOctree() = default;
  APSSOctree(const APSSOctree&) = delete;
  APSSOctree& operator=(const APSSOctree&) = delete;
  ~APSSOctree() { mlod_tree_destroy(m_tree); }

  // APSSOctree::build (CPUUtils.h:176-180); exact-midpoint ties go to tie_policy's child.  With
  // updateRootBoundingBox = false the root keeps its box (CPUUtils.cpp:547): the previous build's,
  // or the one given to setRootBoundingBox.
  template <class V>
  void build(const std::vector<V>& i_points, const std::vector<V>& normals, bool updateRootBoundingBox = true,
             unsigned maxDepth = 12, unsigned maxNumberOfPointsPerLeaf = 1, int tie_policy = 0) {
    if (i_points.size() != normals.size()) throw Error(MLOD_ERR_INVALID, "points and normals differ in length");
    mlod_build_opts o{maxDepth, maxNumberOfPointsPerLeaf, tie_policy, 0u, {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}};
    if (!updateRootBoundingBox) {
      if (!m_haveBox)
        throw Error(MLOD_ERR_STATE, "updateRootBoundingBox = false needs a root box (setRootBoundingBox or a previous build)");
      o.use_root_box = 1u;
      for (int a = 0; a < 6; ++a) o.root_box[a] = m_box[a];
    }
    mlod_tree_destroy(m_tree);
    m_tree = nullptr;
    check(mlod_tree_create(fptr(i_points.data()), fptr(normals.data()), (uint32_t)i_points.size(), &o, nullptr, &m_tree));
    const mlod_tree_info i = info();
    for (int a = 0; a < 6; ++a) m_box[a] = i.bbox[a];
    m_haveBox = true;
  }
  // The root box a build with updateRootBoundingBox = false keeps (octreeNode::setBoundingBox,
  // CPUUtils.h:118).
  void setRootBoundingBox(const float bb[3], const float BB[3]) {
    for (int a = 0; a < 3; ++a) m_box[a] = bb[a], m_box[3 + a] = BB[a];
    m_haveBox = true;
  }

  mlod_tree_info info() const {
    mlod_tree_info i{};
    check(mlod_tree_get_info(m_tree, &i));
    return i;
  }
  float radius() const { return info().radius; }
  unsigned getNbOfLeaves() const { return info().n_points; }

  float getScalingProtectionSphere() const { return m_scalingProtectionSphere; }
  void setScalingProtectionSphere(float v) { m_scalingProtectionSphere = v; }
  unsigned getMinDepth() const { return m_minDepth; }
  void setMinDepth(unsigned v) { m_minDepth = v; }

  // APSSOctree::knnMode / ballMode (CPUUtils.cpp:674-688).
  void knnMode(bool setMode, unsigned knn = 20) {
    m_knnMode = setMode;
    if (m_knnMode) m_ballMode = false;
    m_knn = knn;
  }
  void ballMode(bool setMode, float ballRadius = 0.1f) {
    m_ballMode = setMode;
    if (m_ballMode) m_knnMode = false;
    m_ballRadius = ballRadius;
  }
  bool isAPSS() const { return m_knnMode || m_ballMode; }

  mlod_tree* handle() const {
    if (!m_tree) throw Error(MLOD_ERR_STATE, "octree not built");
    return m_tree;
  }

 private:
  friend class APSS;
  mlod_tree* m_tree = nullptr;
  float m_box[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  bool m_haveBox = false;
  float m_scalingProtectionSphere = 2.0f;
  unsigned m_minDepth = 0;
  bool m_knnMode = false;
  unsigned m_knn = 20;
  bool m_ballMode = false;
  float m_ballRadius = 0.1f;
};

// The projection plugin interface (GPUUtils.h:22-67).
class APSS {
 public:
  virtual ~APSS() { erasePointsFromGPU(); }

  // APSS::project (GPUUtils.h:26-27): n_iterations steps in place on the uploaded queries;
  // `kernel` is ignored as in Fastapss::project (the installed kernel is used).
  virtual void project(unsigned nbOfVectors, float* outputPoints, float* outputNormals, unsigned n_iterations,
                       const Kernel& kernel, float stepMaxSize = -1.f) const {
    (void)kernel;
    sync();
    check(mlod_project(m_apssOctree->handle(), nbOfVectors, n_iterations, stepMaxSize, outputPoints, outputNormals, m_stream));
  }
  template <class V>
  void project(unsigned nbOfVectors, V* outputPoints, V* outputNormals, unsigned n_iterations, const Kernel& kernel,
               float stepMaxSize = -1.f) const {
    project(nbOfVectors, outputPoints ? fptr(outputPoints) : nullptr, outputNormals ? fptr(outputNormals) : nullptr,
            n_iterations, kernel, stepMaxSize);
  }

  // APSS::projectCPU (GPUUtils.h:33-34): the reference's CPU function (REF_CPU semantics,
  // APSSOctree::projectCPU, CPUUtils.cpp:621-636) for one point, computed on the GPU without
  // touching the uploaded queries.  Batch callers (scene::projectPointsOnCPU) should pass all
  // their points to projectPoints instead of looping over this.
  template <class V>
  void projectCPU(const V& qStart, V& outputPoint, V& outputNormal, unsigned n_iterations,
                  const Kernel& kernel = Kernel(), float stepMaxSize = -1.f) const {
    projectPoints(1, &qStart, &outputPoint, &outputNormal, n_iterations, kernel, stepMaxSize, MLOD_SEM_REF_CPU);
  }
  template <class V>
  void projectPoints(unsigned n, const V* points, V* outputPoints, V* outputNormals, unsigned n_iterations,
                     const Kernel& kernel, float stepMaxSize = -1.f, int semantics = -1) const {
    (void)kernel;
    sync(semantics < 0 ? m_semantics : semantics);
    check(mlod_project_points(m_apssOctree->handle(), fptr(points), n, n_iterations, stepMaxSize,
                              outputPoints ? fptr(outputPoints) : nullptr, outputNormals ? fptr(outputNormals) : nullptr,
                              m_stream));
  }

  // APSS::reorganizePoints (GPUUtils.cu:436-450)
  void reorganizePoints(const std::vector<unsigned>* invalidPts, unsigned endValid) {
    static_assert(sizeof(unsigned) == sizeof(uint32_t), "unsigned is 32-bit");
    check(mlod_queries_reorganize(m_apssOctree->handle(), reinterpret_cast<const uint32_t*>(invalidPts->data()),
                                  (uint32_t)invalidPts->size(), endValid, m_stream));
  }

  // APSS::printApss (GPUUtils.cu:426-434)
  void printApss() const {
    const mlod_tree_info i = m_apssOctree->info();
    std::printf("m_numberOfNodes: %u\nm_scalingProtectionSphere: %g\nm_minDepth: %u\n", i.n_nodes,
                (double)m_scalingProtectionSphere, m_minDepth);
  }
  // APSS::stop (GPUUtils.cu:394-397) stops the CUDA profiler; there is nothing to stop here.
  void stop() {}
  // APSS::eraseAndRestart (GPUUtils.cu:399-404): like the reference, the APSS owns and deletes
  // its octree here and continues with a new, empty one.
  void eraseAndRestart() {
    eraseFromGPU();
    delete m_apssOctree;
    m_apssOctree = new APSSOctree();
  }

  // Fastapss::copyToGPU (Fastapss.cu:457-484): the tree is already resident; installs the kernel.
  virtual void copyToGPU(const Kernel& kernel, unsigned nbOfVectors) {
    (void)nbOfVectors;
    updateKernel(kernel);
  }
  virtual void eraseFromGPU() { erasePointsFromGPU(); }

  // APSS::copyPointsToGPU / erasePointsFromGPU (GPUUtils.cu:406-411, 452-464).
  virtual void copyPointsToGPU(unsigned nbOfVectors, const float* points) {
    check(mlod_queries_upload(m_apssOctree->handle(), points, nbOfVectors, m_stream));
  }
  template <class V>
  void copyPointsToGPU(unsigned nbOfVectors, const V* points) {
    copyPointsToGPU(nbOfVectors, fptr(points));
  }
  virtual void erasePointsFromGPU() {
    if (m_apssOctree && m_apssOctree->m_tree) mlod_queries_erase(m_apssOctree->m_tree);
  }

  // APSS::updateKernel (GPUUtils.cu:466-472).
  void updateKernel(const Kernel& kernel) {
    const mlod_kernel k = kernel.to_c();
    check(mlod_set_kernel(m_apssOctree->handle(), &k));
  }

  void setScalingProtectionSphere(float val) {
    m_scalingProtectionSphere = val;
    m_apssOctree->setScalingProtectionSphere(val);
  }
  float getScalingProtectionSphere() const { return m_scalingProtectionSphere; }
  void setMinDepth(unsigned val) {
    m_minDepth = val;
    m_apssOctree->setMinDepth(val);
  }
  unsigned getMinDepth() const { return m_minDepth; }

  void set_semantics(int semantics) { m_semantics = semantics; }
  void set_precision(int precision) { m_precision = precision; }
  void set_traversal(int traversal) { m_traversal = traversal; }
  void set_stream(void* hip_stream) { m_stream = hip_stream; }
  const APSSOctree* CPU() const { return m_apssOctree; }

 protected:
  friend class OctreeGenerator;
  friend class DualContouring;
  APSS(APSSOctree* octree, int mode) : m_apssOctree(octree), m_mode(mode) {
    m_scalingProtectionSphere = octree->getScalingProtectionSphere();
    m_minDepth = octree->getMinDepth();
  }
  void sync(int semantics = -1) const {
    mlod_tree* t = m_apssOctree->handle();
    check(mlod_set_params(t, m_scalingProtectionSphere, m_minDepth, semantics < 0 ? m_semantics : semantics, m_precision));
    int mode = m_mode;
    if (mode == MLOD_MODE_MLOD && m_apssOctree->m_knnMode) mode = MLOD_MODE_KNN;
    if (mode == MLOD_MODE_MLOD && m_apssOctree->m_ballMode) mode = MLOD_MODE_BALL;
    check(mlod_set_mode(t, mode, m_apssOctree->m_knn, m_apssOctree->m_ballRadius));
    check(mlod_set_traversal(t, m_traversal));
  }

  APSSOctree* m_apssOctree;
  int m_mode;
  float m_scalingProtectionSphere = 2.f;
  unsigned m_minDepth = 0;
  int m_semantics = MLOD_SEM_REF_GPU;
  int m_precision = MLOD_PREC_FP32;
  int m_traversal = MLOD_TRAVERSAL_LANE;
  void* m_stream = nullptr;
};

// MLoD projection (Fastapss, Fastapss.h:32-51).
class Fastapss : public APSS {
 public:
  explicit Fastapss(APSSOctree* apssOctree) : APSS(apssOctree, MLOD_MODE_MLOD) {}
  Fastapss(APSSOctree* apssOctree, const Kernel& kernel, unsigned nbOfVectors) : Fastapss(apssOctree) {
    copyToGPU(kernel, nbOfVectors);
  }
};

// Brute-force APSS over all points (GlobalAPSS, GlobalAPSS.h:19-43).
class GlobalAPSS : public APSS {
 public:
  explicit GlobalAPSS(APSSOctree* apssOctree) : APSS(apssOctree, MLOD_MODE_GLOBAL) {}
  GlobalAPSS(APSSOctree* apssOctree, const Kernel& kernel, unsigned nbOfVectors) : GlobalAPSS(apssOctree) {
    copyToGPU(kernel, nbOfVectors);
  }
};

// octree_id (sources/dual-contouring/octree.h:18-44): a cell of the [-1, 1]^3 octree.
struct octree_id {
  int depth;
  int pos[3];
};

// OctreeGenerator (sources/scene/upsample.h:26-52), run on the GPU.  generate projects with
// apss's tree, kernel and settings, as find_stable_nodesGPU does (upsample.cpp:317-346); the
// cells come back in the reference's node_comparator order.
class OctreeGenerator {
 public:
  OctreeGenerator(APSS* apss, const Kernel& kernel) : apss(apss), kernel(kernel) {
    check(mlod_generator_create(apss ? apss->m_apssOctree->handle() : nullptr, &m_gen));
  }
  OctreeGenerator(const OctreeGenerator&) = delete;
  OctreeGenerator& operator=(const OctreeGenerator&) = delete;
  ~OctreeGenerator() { mlod_generator_destroy(m_gen); }

  // OctreeGenerator::generate (upsample.cpp:228-251)
  std::vector<octree_id> generate(unsigned int max_depth, std::vector<octree_id>* total_nodes = nullptr) const {
    if (!apss) throw Error(MLOD_ERR_STATE, "generate needs an APSS");
    apss->sync();
    check(mlod_generate(m_gen, max_depth, total_nodes != nullptr, apss->m_stream));
    if (total_nodes) *total_nodes = cells(1);
    return cells(0);
  }

  // OctreeGenerator::generateFromPoints (upsample.cpp:128-144); box6 = (min xyz, max xyz),
  // NULL for [-1, 1]^3.
  template <class V>
  std::vector<octree_id> generateFromPoints(unsigned int max_depth, const std::vector<V>& pts,
                                            const float* box6 = nullptr) const {
    check(mlod_generate_from_points(m_gen, fptr(pts.data()), (uint32_t)pts.size(), 0, max_depth, box6, nullptr));
    return cells(0);
  }

  // Centres of the generated cells (aabb_octree::center): the upsampling query points.
  std::vector<float> centres() const {
    uint64_t inf[6];
    check(mlod_generator_info(m_gen, inf));
    std::vector<float> out(3 * inf[0]);
    check(mlod_generator_centres(m_gen, out.data(), nullptr, nullptr));
    return out;
  }

 private:
  std::vector<octree_id> cells(int which) const {
    uint64_t inf[6];
    check(mlod_generator_info(m_gen, inf));
    std::vector<octree_id> out(which ? inf[1] : inf[0]);
    static_assert(sizeof(octree_id) == 4 * sizeof(int32_t), "octree_id is four int32");
    if (!out.empty()) check(mlod_generator_cells(m_gen, which, reinterpret_cast<int32_t*>(out.data()), nullptr));
    return out;
  }
  APSS* apss;
  const Kernel& kernel;
  mlod_generator* m_gen = nullptr;
};

}  // namespace mlod

#endif  // MLOD_APSS_HPP

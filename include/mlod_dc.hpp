// mlod_dc.hpp — header-only C++ for the reference's mesh output from the projection, on
// libmlod.so (mlod.h): the dual contouring of the adaptive octree and the mesh passes of
// scene::computeDualContouring (sources/scene/scene.cpp:1252-1452), with the reference's names.
//
// * mlod::DualContouring — uniformize (upsample.cpp:262-272) of octree cells, the corner field
//   by one projection step of every leaf corner (scene.cpp:1346-1387), and
//   dual_contouring::extract (dual_contouring.h:59-87).  Vertices come back in the projection
//   frame (new2bbox applied), triangles as vertex index triples.
// * mlod::mesh::smoothMesh / subdivideMesh (upsample.h:55-133), the reference's results bit for
//   bit, on the GPU.
// * mlod::computeDualContouring — the reference's whole sequence on the adaptive octree:
//   OctreeGenerator::generate with its crossed cells, uniformize, corners, extraction, the
//   vertices projected 3 steps, then twice smoothing, subdivision and projection;
//   computeDualContouringOnPoints — the same on an octree built on points (uniformizeFromPoints).
// * mlod::meshComments — the kernel and octree comment lines the reference writes into the mesh
//   file (scene.cpp:1498-1530); Pn::saveMesh (mlod_pn.hpp) writes the file.
//
// Vertices and triangles are the reference's; vertices are numbered in octree node order where
// the reference numbers them in its traversal's order.
#ifndef MLOD_DC_HPP
#define MLOD_DC_HPP

#include <cstdint>
#include <sstream>
#include <string>
#include <vector>

#include "mlod_apss.hpp"
#include "mlod_pn.hpp"

namespace mlod {

class DualContouring {
 public:
  // uniformize(max_depth, cells, ..., cut) of the cells (usually OctreeGenerator's crossed cells),
  // or with fromPoints uniformizeFromPoints(cells) (upsample.cpp:253-259: cells of an octree
  // built on points by OctreeGenerator::generateFromPoints; no re-rooting, no cut)
  DualContouring(const std::vector<octree_id>& cells, unsigned max_depth = 0, bool cut = false, void* stream = nullptr,
                 bool fromPoints = false)
      : m_stream(stream) {
    check(mlod_dc_create(&m_dc));
    static_assert(sizeof(octree_id) == 4 * sizeof(int32_t), "octree_id is four int32");
    const int32_t* ids = reinterpret_cast<const int32_t*>(cells.data());
    const int r = fromPoints ? mlod_dc_build_from_points(m_dc, ids, (uint32_t)cells.size(), m_stream)
                             : mlod_dc_build(m_dc, ids, (uint32_t)cells.size(), max_depth, cut ? 1 : 0, m_stream);
    if (r != MLOD_OK) {
      mlod_dc_destroy(m_dc);
      check(r);
    }
  }
  DualContouring(const DualContouring&) = delete;
  DualContouring& operator=(const DualContouring&) = delete;
  ~DualContouring() { mlod_dc_destroy(m_dc); }

  // [nodes, leaves, corners, finest depth, diag, origin x y z, extracted]
  std::vector<double> info() const {
    std::vector<double> o(9);
    check(mlod_dc_info(m_dc, o.data()));
    return o;
  }

  // the corner field from one projection step of every leaf corner with apss's tree and settings
  void projectCorners(const APSS& apss) {
    apss.sync();
    check(mlod_dc_project_corners(m_dc, apss.m_apssOctree->handle(), m_stream));
  }
  // or the caller's values, one per corner in the order of the library's corner list
  void setCornerValues(const std::vector<float>& f) { check(mlod_dc_set_corner_values(m_dc, f.data(), m_stream)); }

  // the extracted mesh refined on the device (scene.cpp:1417-1452): vertices projected `steps`
  // steps with apss's tree and settings, then `rounds` times smoothing, subdivision, projection
  void refine(const APSS& apss, unsigned steps, unsigned rounds, std::vector<Vec3f>& vertices,
              std::vector<Vec3f>& normals, std::vector<uint32_t>& indices) {
    apss.sync();
    uint32_t nv = 0, nt = 0;
    check(mlod_dc_refine(m_dc, apss.m_apssOctree->handle(), steps, rounds, &nv, &nt, m_stream));
    vertices.assign(nv, Vec3f{});
    normals.assign(nv, Vec3f{});
    indices.assign(3 * (size_t)nt, 0);
    check(mlod_dc_refined_mesh(m_dc, nv ? &vertices[0].x : nullptr, nv ? &normals[0].x : nullptr,
                               nt ? indices.data() : nullptr, m_stream));
  }

  // dual_contouring::extract: vertices in the projection frame, triangles (3 indices each)
  void extract(std::vector<Vec3f>& vertices, std::vector<uint32_t>& indices) {
    uint32_t nv = 0, nt = 0;
    check(mlod_dc_extract(m_dc, &nv, &nt, m_stream));
    vertices.assign(nv, Vec3f{});
    indices.assign(3 * (size_t)nt, 0);
    check(mlod_dc_mesh(m_dc, nv ? &vertices[0].x : nullptr, nullptr, nullptr, nt ? indices.data() : nullptr, m_stream));
  }

 private:
  mlod_dc* m_dc = nullptr;
  void* m_stream = nullptr;
};

namespace mesh {

// mesh::smoothMesh (upsample.h:55-76)
inline void smoothMesh(std::vector<Vec3f>& iPoints, const std::vector<uint32_t>& iTris) {
  std::vector<Vec3f> out(iPoints.size());
  if (iPoints.empty()) return;
  check(mlod_mesh_smooth(&iPoints[0].x, (uint32_t)iPoints.size(), iTris.data(), (uint32_t)(iTris.size() / 3),
                         &out[0].x, nullptr));
  iPoints.swap(out);
}

// mesh::subdivideMesh (upsample.h:78-133)
inline void subdivideMesh(std::vector<Vec3f>& iPoints, std::vector<uint32_t>& iTris) {
  const uint32_t nv = (uint32_t)iPoints.size(), nt = (uint32_t)(iTris.size() / 3);
  std::vector<Vec3f> out((size_t)nv + 3 * (size_t)nt);
  std::vector<uint32_t> tris(12 * (size_t)nt);
  uint32_t nOut = 0;
  if (out.empty()) return;
  check(mlod_mesh_subdivide(nv ? &iPoints[0].x : &out[0].x, nv, iTris.data(), nt, &out[0].x, (uint32_t)out.size(),
                            tris.data(), &nOut, nullptr));
  out.resize(nOut);
  iPoints.swap(out);
  iTris.swap(tris);
}

}  // namespace mesh

// scene::computeDualContouring on the adaptive octree (scene.cpp:1252-1452, m_useAdaptiveOctree):
// vertices and normals in the projection frame, triangles as index triples.
inline void computeDualContouring(const APSS& apss, const Kernel& kernel, unsigned maxDepth, bool cut,
                                  std::vector<Vec3f>& vertices, std::vector<Vec3f>& normals,
                                  std::vector<uint32_t>& indices, unsigned reprojectSteps = 3, unsigned refine = 2) {
  std::vector<octree_id> total;
  {
    OctreeGenerator gen(const_cast<APSS*>(&apss), kernel);
    gen.generate(maxDepth, &total);
  }
  DualContouring dc(total, maxDepth, cut);
  dc.projectCorners(apss);
  dc.extract(vertices, indices);
  dc.refine(apss, reprojectSteps, refine, vertices, normals, indices);
}

// scene::computeDualContouring on an octree built on points (scene.cpp:1279-1299: the usual
// octree on the input points, or the octree on the projected points): the cells of
// OctreeGenerator::generateFromPoints(maxDepth, points), uniformizeFromPoints, then as above.
template <class V>
inline void computeDualContouringOnPoints(const APSS& apss, const Kernel& kernel, unsigned maxDepth,
                                          const std::vector<V>& points, std::vector<Vec3f>& vertices,
                                          std::vector<Vec3f>& normals, std::vector<uint32_t>& indices,
                                          unsigned reprojectSteps = 3, unsigned refine = 2) {
  std::vector<octree_id> cells;
  {
    OctreeGenerator gen(const_cast<APSS*>(&apss), kernel);
    cells = gen.generateFromPoints(maxDepth, points);
  }
  DualContouring dc(cells, maxDepth, false, nullptr, true);
  dc.projectCorners(apss);
  dc.extract(vertices, indices);
  dc.refine(apss, reprojectSteps, refine, vertices, normals, indices);
}

// The comment lines of the reference's mesh file (scene.cpp:1498-1530) for a kernel whose sigmas
// were scaled by a cloud of bounding-box radius `radius`, on the adaptive octree.
inline std::vector<std::string> meshComments(const Kernel& k, float radius, bool adaptiveOctree = true) {
  std::vector<std::string> c{"Created by mlodsurfaces_amd", "Kernel parameters:"};
  auto line = [&c](const std::string& head, auto v) {
    std::ostringstream o;
    o << head << v;
    c.push_back(o.str());
  };
  switch (k.type) {
    case Kernel::GAUSSIAN_MULTIPLE:
      c.push_back("Multiple Gaussian");
      line("First sigma  ", k.beginningSigma);
      line("Factor ", k.factor);
      line("Number of sigmas ", k.sigmas.size());
      line("Constant_shift (unused) ", k.constant_shift);
      break;
    case Kernel::SINGULAR:
      c.push_back("Singular");
      line("Exponent ", k.exponent);
      line("Constant_shift ", k.constant_shift);
      break;
    case Kernel::GAUSSIAN:
      c.push_back("Gaussian");
      line("Sigma ", k.sigma / radius);
      line("Constant_shift ", k.constant_shift);
      break;
    default:
      c.push_back("Kernel unknown");
  }
  c.push_back("Octree:");
  c.push_back(adaptiveOctree ? "Built using the adaptive octree method" : "Built on input points");
  return c;
}

}  // namespace mlod

#endif  // MLOD_DC_HPP

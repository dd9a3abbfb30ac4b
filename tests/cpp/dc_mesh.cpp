// Test driver (GPU): dc_mesh <model.ply|pn> <model.params> <out.ply> [points] — the reference's
// surface mesh of a model through the C++ headers: Pn load + normalize, .params,
// APSSOctree::build, Fastapss with the model's kernel, cutoff and min depth,
// computeDualContouring at the model's max depth (with "points": computeDualContouringOnPoints on
// the octree built on the normalised points), and the mesh file in the file frame.
#include "mlod_dc.hpp"

#include <cstdio>
#include <string>

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  try {
    mlod::Pn pn;
    pn.load(argv[1]);
    pn.normalize();
    mlod::ModelParams prm;
    prm.load(argv[2]);
    mlod::APSSOctree tree;
    tree.build(pn.positions, pn.normals);
    mlod::Kernel k = prm.kernel(tree.radius());
    mlod::Fastapss apss(&tree, k, pn.positions.size());
    apss.setScalingProtectionSphere(prm.cutoff);
    apss.setMinDepth(prm.minDepth);
    apss.set_precision(MLOD_PREC_FP32C);  // the compensated form, as the Python command line
    std::vector<mlod::Vec3f> v, n;
    std::vector<uint32_t> tris;
    const bool onPoints = argc > 4 && std::string(argv[4]) == "points";
    if (onPoints)
      mlod::computeDualContouringOnPoints(apss, k, prm.maxDepth, pn.positions, v, n, tris);
    else
      mlod::computeDualContouring(apss, k, prm.maxDepth, false, v, n, tris);
    pn.saveMesh(argv[3], v, n, tris, mlod::meshComments(k, tree.radius(), !onPoints));
    std::printf("mesh of %zu vertices and %zu triangles at depth %u\n", v.size(), tris.size() / 3, prm.maxDepth);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}

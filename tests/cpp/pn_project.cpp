// Test driver (GPU): pn_project <model.ply|pn> <model.params> <out.ply> — the reference's
// automatic projection of a model's own points through the C++ headers: Pn load + normalize,
// .params, APSSOctree::build, Fastapss with the model's kernel, cutoff and min depth,
// maxNbIterations steps, projected points saved in the file frame.
#include "mlod_apss.hpp"
#include "mlod_pn.hpp"

#include <cstdio>

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
    apss.copyPointsToGPU(pn.positions.size(), pn.positions.data());
    std::vector<mlod::Vec3f> out(pn.positions.size()), nrm(out.size());
    apss.project(out.size(), out.data(), nrm.data(), prm.maxNbIterations, k);
    pn.save(argv[3], out, nrm);
    std::printf("projected %zu points, %zu sigmas, %u steps\n", out.size(), k.sigmas.size(), prm.maxNbIterations);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}

// Test driver for include/mlod_pn.hpp (no GPU): pn_tool <in.ply|pn> <out_prefix> [<model.params>]
// writes <out_prefix>.norm (float32: offset[3], norme, then the normalised positions), saves the
// normalised set back to <out_prefix>.ply (file frame), the normalised positions in Pn::order's
// Morton order to <out_prefix>.order, and prints the .params fields and the
// kernel's sigmas for radius 0.8125; with the .params, also writes <out_prefix>Mesh.ply, a mesh
// of the first four points (two triangles) with the mesh-file comments of that kernel.
#include "mlod_dc.hpp"

#include <cstdio>

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  mlod::Pn pn;
  pn.load(argv[1]);
  pn.normalize();
  const std::string out = argv[2];
  FILE* f = std::fopen((out + ".norm").c_str(), "wb");
  std::fwrite(&pn.offset, sizeof(float), 3, f);
  std::fwrite(&pn.norme, sizeof(float), 1, f);
  std::fwrite(pn.positions.data(), sizeof(mlod::Vec3f), pn.positions.size(), f);
  std::fclose(f);
  pn.save(out + ".ply");
  {  // Pn::order of the normalised positions
    mlod::Pn o = pn;
    o.order();
    FILE* g = std::fopen((out + ".order").c_str(), "wb");
    std::fwrite(o.positions.data(), sizeof(mlod::Vec3f), o.positions.size(), g);
    std::fclose(g);
  }
  std::printf("points %zu normals %zu\n", pn.positions.size(), pn.normals.size());
  if (argc > 3) {
    mlod::ModelParams p;
    p.load(argv[3]);
    std::printf("type %d cutoff %.9g min_depth %u iterations %u computed %d n_sigmas %u factor %.9g begin %.9g\n",
                p.kernelType, p.cutoff, p.minDepth, p.maxNbIterations, (int)p.computedSigma, p.nbOfComputedSigmas,
                p.factor, p.beginningSigma);
    const mlod::Kernel k = p.kernel(0.8125f);
    std::printf("sigma %.9g sigmas", k.sigma);
    for (float s : k.sigmas) std::printf(" %.9g", s);
    std::printf("\n");
    if (pn.positions.size() >= 4) {
      const std::vector<mlod::Vec3f> v(pn.positions.begin(), pn.positions.begin() + 4);
      const std::vector<mlod::Vec3f> n(pn.normals.begin(), pn.normals.begin() + 4);
      pn.saveMesh(out + "Mesh.ply", v, n, {0, 1, 2, 0, 2, 3}, mlod::meshComments(k, 0.8125f));
    }
  }
  return 0;
}

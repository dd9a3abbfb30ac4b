// Drop-in usage of include/mlod_apss.hpp the way the reference's scene code drives its
// APSS plugin (scene.cpp:916-966): build the octree, install the kernel with sigmas scaled by
// the bbox radius (scene.cpp:1025-1033), upload queries, project one step at a time.
// Reads points/normals/queries as raw float32 files, writes projected points and normals, and,
// given a sixth path, the cells of the adaptive octree (OctreeGenerator::generate at depth 4,
// as scene::generate_adaptive_octree drives it, scene.cpp:155-167) as int32 rows.
#include <cstdio>
#include <fstream>
#include <vector>

#include "mlod_apss.hpp"

struct vec3 {
  float x, y, z;
};

static std::vector<vec3> read(const char* path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  const size_t n = (size_t)f.tellg() / sizeof(vec3);
  std::vector<vec3> v(n);
  f.seekg(0);
  f.read(reinterpret_cast<char*>(v.data()), n * sizeof(vec3));
  return v;
}

int main(int argc, char** argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s points normals queries out_points out_normals\n", argv[0]);
    return 2;
  }
  const std::vector<vec3> P = read(argv[1]), N = read(argv[2]), Q = read(argv[3]);
  try {
    mlod::APSSOctree octree;
    octree.build(P, N);  // maxDepth 12, 1 point per leaf
    mlod::Kernel k;
    k.type = mlod::Kernel::GAUSSIAN_MULTIPLE;
    k.beginningSigma = 0.01f;
    k.factor = 2.f;
    k.nbOfComputedSigmas = 4;
    k.updateSigmas(true, octree.radius());
    mlod::Fastapss apss(&octree, k, (unsigned)Q.size());
    apss.setScalingProtectionSphere(3.f);
    apss.copyPointsToGPU((unsigned)Q.size(), Q.data());
    std::vector<vec3> outP(Q.size()), outN(Q.size());
    for (int it = 0; it < 3; ++it) apss.project((unsigned)Q.size(), outP.data(), outN.data(), 1, k);
    std::ofstream(argv[4], std::ios::binary).write(reinterpret_cast<const char*>(outP.data()), outP.size() * sizeof(vec3));
    std::ofstream(argv[5], std::ios::binary).write(reinterpret_cast<const char*>(outN.data()), outN.size() * sizeof(vec3));
    if (argc > 6) {
      mlod::OctreeGenerator gen(&apss, k);
      std::vector<mlod::octree_id> total;
      const std::vector<mlod::octree_id> cells = gen.generate(4, &total);
      std::ofstream(argv[6], std::ios::binary)
          .write(reinterpret_cast<const char*>(cells.data()), cells.size() * sizeof(mlod::octree_id));
      std::printf("cells %zu total %zu\n", cells.size(), total.size());
    }
  } catch (const mlod::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}

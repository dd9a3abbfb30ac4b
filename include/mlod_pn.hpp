// mlod_pn.hpp — header-only C++ point-set files and per-model settings of the reference
// application, for callers of the projection (no GPU, no libmlod.so needed).
//
// * mlod::Pn — an oriented point set as the reference's `Pn` loads, normalises and saves it
//   (sources/apss/Pn.cpp:102-164 load, 201-245 save, 262-290 normalize): PN files (float32
//   records x y z nx ny nz) and PLY files (binary little-endian or ASCII; vertex properties
//   x y z [nx ny nz] of any scalar type; later elements ignored).  save() writes the reference's
//   PLY header and records with positions mapped back as p * norme + offset, as Pn::save and
//   scene::saveProjectedPoints (sources/scene/scene.cpp:1622-1661) do; saveMesh() writes the
//   surface mesh file of scene::computeDualContouring (scene.cpp:1489-1560).
// * mlod::ModelParams — the 134-byte .params file of a model (scene::loadParameters,
//   scene.cpp:1183-1250), and kernel(radius): the mlod::Kernel it describes, sigmas generated
//   and scaled as Kernel::updateSigmas (Kernel.cpp:110-128).  Include mlod_apss.hpp first to get
//   kernel(); the rest needs only the standard library.
#ifndef MLOD_PN_HPP
#define MLOD_PN_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mlod {

struct Vec3f {
  float x = 0.f, y = 0.f, z = 0.f;
};

class Pn {
 public:
  std::vector<Vec3f> positions, normals;  // normals empty when the file has none
  Vec3f offset;                           // file position = p * norme + offset
  float norme = 1.f;

  static bool ends_with(const std::string& s, const std::string& e) {
    return s.size() >= e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0;
  }

  // Pn::load without the optional normalisation (call normalize() for it).
  void load(const std::string& path) {
    positions.clear();
    normals.clear();
    offset = Vec3f{};
    norme = 1.f;
    if (ends_with(path, "pn")) {
      std::ifstream f(path, std::ios::binary);
      if (!f) throw std::runtime_error(path + " is not a valid PN file");
      float r[6];
      while (f.read(reinterpret_cast<char*>(r), sizeof r)) {
        positions.push_back({r[0], r[1], r[2]});
        normals.push_back({r[3], r[4], r[5]});
      }
    } else if (ends_with(path, "ply")) {
      load_ply(path);
    } else {
      throw std::runtime_error(path + ": neither .ply nor .pn");
    }
  }

  // Pn::normalize (Pn.cpp:262-290), float arithmetic in the same order.
  void normalize() {
    if (positions.empty()) throw std::runtime_error("empty point set");
    Vec3f off = positions[0], mx = positions[0], mn = positions[0];
    for (size_t i = 1; i < positions.size(); ++i) {
      const Vec3f& p = positions[i];
      off.x += p.x, off.y += p.y, off.z += p.z;
      mx.x = std::max(p.x, mx.x), mx.y = std::max(p.y, mx.y), mx.z = std::max(p.z, mx.z);
      mn.x = std::min(p.x, mn.x), mn.y = std::min(p.y, mn.y), mn.z = std::min(p.z, mn.z);
    }
    const float n = static_cast<float>(positions.size());
    off.x /= n, off.y /= n, off.z /= n;
    mx.x -= off.x, mx.y -= off.y, mx.z -= off.z;
    mn.x -= off.x, mn.y -= off.y, mn.z -= off.z;
    const float dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
    // glm::length: (x*x + y*y) + z*z, each product rounded (volatile keeps them unfused)
    const volatile float sx = dx * dx, sy = dy * dy, sz = dz * dz;
    const float small = std::sqrt((sx + sy) + sz) * 0.005f;
    mx.x += small, mx.y += small, mx.z += small;
    mn.x -= small, mn.y -= small, mn.z -= small;
    norme = std::max({std::fabs(mx.x), std::fabs(mx.y), std::fabs(mx.z), std::fabs(mn.x), std::fabs(mn.y),
                      std::fabs(mn.z)});
    offset = off;
    for (Vec3f& p : positions) p = {(p.x - off.x) / norme, (p.y - off.y) / norme, (p.z - off.z) / norme};
  }

  // Pn::order (Pn.cpp:388-445): the positions sorted along the Morton curve of a sizeOfGrid^3
  // grid over the box [mini, maxi]^3 (mini / maxi: smallest / largest coordinate of the current
  // positions over all axes, as computeMinMax leaves them for the query sets the reference
  // orders).  Points of equal code keep their order (the reference's std::sort leaves it
  // unspecified); like the reference, only the positions are reordered.
  void order(unsigned sizeOfGrid = 64) {
    if (positions.empty()) return;
    Vec3f lo = positions[0], hi = positions[0];
    for (const Vec3f& p : positions) {
      lo.x = std::min(lo.x, p.x), lo.y = std::min(lo.y, p.y), lo.z = std::min(lo.z, p.z);
      hi.x = std::max(hi.x, p.x), hi.y = std::max(hi.y, p.y), hi.z = std::max(hi.z, p.z);
    }
    const float mini = std::min(std::min(lo.x, lo.y), lo.z), maxi = std::max(std::max(hi.x, hi.y), hi.z);
    const float norme = maxi - mini;
    auto cell = [&](float v) {
      const float c = std::floor((v - mini) / norme * (float)sizeOfGrid);
      return c <= 0.f ? 0u : std::min((uint32_t)c, sizeOfGrid - 1u);
    };
    auto deposit = [](uint32_t v, uint32_t mask) {  // _pdep_u32
      uint32_t r = 0;
      for (uint32_t bit = 1; mask; bit <<= 1) {
        const uint32_t low = mask & (~mask + 1u);
        if (v & bit) r |= low;
        mask &= mask - 1u;
      }
      return r;
    };
    std::vector<std::pair<uint32_t, uint32_t>> code(positions.size());
    for (size_t i = 0; i < positions.size(); ++i)
      code[i] = {deposit(cell(positions[i].x), 0x92492492u) | deposit(cell(positions[i].y), 0x49249249u) |
                     deposit(cell(positions[i].z), 0x24924924u),
                 (uint32_t)i};
    std::stable_sort(code.begin(), code.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<Vec3f> sorted(positions.size());
    for (size_t i = 0; i < code.size(); ++i) sorted[i] = positions[code[i].second];
    positions.swap(sorted);
  }

  // Pn::save: PN records or a binary little-endian PLY, positions mapped back to the file frame.
  void save(const std::string& path) const { save(path, positions, normals); }
  void save(const std::string& path, const std::vector<Vec3f>& pts, const std::vector<Vec3f>& nrm) const {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    const bool ply = ends_with(path, "ply");
    if (!ply && !ends_with(path, "pn")) throw std::runtime_error(path + ": neither .ply nor .pn");
    if (ply) {
      f << "ply\nformat binary_little_endian 1.0\nelement vertex " << pts.size()
        << "\nproperty float x\nproperty float y\nproperty float z\n";
      if (!nrm.empty()) f << "property float nx\nproperty float ny\nproperty float nz\n";
      f << "end_header\n";
    }
    for (size_t i = 0; i < pts.size(); ++i) {
      const Vec3f p{pts[i].x * norme + offset.x, pts[i].y * norme + offset.y, pts[i].z * norme + offset.z};
      f.write(reinterpret_cast<const char*>(&p), sizeof p);
      if (!nrm.empty()) f.write(reinterpret_cast<const char*>(&nrm[i]), sizeof(Vec3f));
      else if (!ply) f.write("\0\0\0\0\0\0\0\0\0\0\0\0", sizeof(Vec3f));
    }
  }

  // The surface mesh file of scene::computeDualContouring (scene.cpp:1489-1560): binary
  // little-endian PLY, "comment" lines, vertices x y z nx ny nz with positions mapped back to the
  // file frame, faces as a uchar 3 and three int indices.
  void saveMesh(const std::string& path, const std::vector<Vec3f>& verts, const std::vector<Vec3f>& nrm,
                const std::vector<uint32_t>& indices, const std::vector<std::string>& comments) const {
    if (nrm.size() != verts.size()) throw std::runtime_error("saveMesh: one normal per vertex expected");
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f << "ply\nformat binary_little_endian 1.0\n";
    for (const std::string& c : comments) f << "comment " << c << "\n";
    f << "element vertex " << verts.size()
      << "\nproperty float x\nproperty float y\nproperty float z\nproperty float nx\nproperty float ny\n"
         "property float nz\nelement face "
      << indices.size() / 3 << "\nproperty list uchar int vertex_indices\nend_header\n";
    for (size_t i = 0; i < verts.size(); ++i) {
      const Vec3f p{verts[i].x * norme + offset.x, verts[i].y * norme + offset.y, verts[i].z * norme + offset.z};
      f.write(reinterpret_cast<const char*>(&p), sizeof p);
      f.write(reinterpret_cast<const char*>(&nrm[i]), sizeof(Vec3f));
    }
    for (size_t t = 0; t + 2 < indices.size(); t += 3) {
      const char three = 3;
      f.write(&three, 1);
      f.write(reinterpret_cast<const char*>(&indices[t]), 3 * sizeof(uint32_t));
    }
  }

 private:
  struct Prop {
    std::string name, type;
    bool list = false;
    std::string countType;
  };
  struct Element {
    std::string name;
    size_t count = 0;
    std::vector<Prop> props;
  };

  static size_t type_size(const std::string& t) {
    if (t == "char" || t == "int8" || t == "uchar" || t == "uint8") return 1;
    if (t == "short" || t == "int16" || t == "ushort" || t == "uint16") return 2;
    if (t == "int" || t == "int32" || t == "uint" || t == "uint32" || t == "float" || t == "float32") return 4;
    if (t == "double" || t == "float64") return 8;
    throw std::runtime_error("unknown PLY type " + t);
  }
  static double read_scalar(const char* p, const std::string& t) {
    if (t == "float" || t == "float32") { float v; std::memcpy(&v, p, 4); return v; }
    if (t == "double" || t == "float64") { double v; std::memcpy(&v, p, 8); return v; }
    if (t == "char" || t == "int8") return static_cast<int8_t>(*p);
    if (t == "uchar" || t == "uint8") return static_cast<uint8_t>(*p);
    if (t == "short" || t == "int16") { int16_t v; std::memcpy(&v, p, 2); return v; }
    if (t == "ushort" || t == "uint16") { uint16_t v; std::memcpy(&v, p, 2); return v; }
    if (t == "int" || t == "int32") { int32_t v; std::memcpy(&v, p, 4); return v; }
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
  }

  void load_ply(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(path + " is not a valid ply file");
    std::string line, format;
    std::getline(f, line);
    if (line.compare(0, 3, "ply") != 0) throw std::runtime_error(path + ": not a PLY file");
    std::vector<Element> els;
    while (std::getline(f, line)) {
      std::istringstream ss(line);
      std::string w;
      ss >> w;
      if (w == "format") {
        ss >> format;
      } else if (w == "element") {
        Element e;
        ss >> e.name >> e.count;
        els.push_back(e);
      } else if (w == "property") {
        if (els.empty()) throw std::runtime_error("PLY property before any element");
        Prop p;
        ss >> p.type;
        if (p.type == "list") {
          p.list = true;
          ss >> p.countType >> p.type;
        }
        ss >> p.name;
        els.back().props.push_back(p);
      } else if (w == "end_header") {
        break;
      }
    }
    if (format != "binary_little_endian" && format != "ascii")
      throw std::runtime_error(path + ": unsupported PLY format " + format);
    for (const Element& e : els) {
      if (e.name != "vertex") {
        if (format == "ascii") {  // skip the element's lines
          for (size_t i = 0; i < e.count; ++i) std::getline(f, line);
          continue;
        }
        for (size_t i = 0; i < e.count; ++i)
          for (const Prop& p : e.props) {
            if (p.list) {
              std::vector<char> c(type_size(p.countType));
              f.read(c.data(), c.size());
              const size_t k = static_cast<size_t>(read_scalar(c.data(), p.countType));
              f.seekg(static_cast<std::streamoff>(k * type_size(p.type)), std::ios::cur);
            } else {
              f.seekg(static_cast<std::streamoff>(type_size(p.type)), std::ios::cur);
            }
          }
        continue;
      }
      int ix[6] = {-1, -1, -1, -1, -1, -1};
      const char* names[6] = {"x", "y", "z", "nx", "ny", "nz"};
      std::vector<size_t> off(e.props.size());
      size_t rec = 0;
      for (size_t j = 0; j < e.props.size(); ++j) {
        if (e.props[j].list) throw std::runtime_error("list properties in the vertex element are not supported");
        off[j] = rec;
        rec += type_size(e.props[j].type);
        for (int c = 0; c < 6; ++c)
          if (e.props[j].name == names[c]) ix[c] = static_cast<int>(j);
      }
      if (ix[0] < 0 || ix[1] < 0 || ix[2] < 0) throw std::runtime_error(path + ": PLY file without vertex x y z");
      const bool hasN = ix[3] >= 0 && ix[4] >= 0 && ix[5] >= 0;
      positions.resize(e.count);
      if (hasN) normals.resize(e.count);
      std::vector<char> buf(rec);
      std::vector<double> v(e.props.size());
      for (size_t i = 0; i < e.count; ++i) {
        if (format == "ascii") {
          for (double& x : v) f >> x;
        } else {
          f.read(buf.data(), rec);
          for (size_t j = 0; j < e.props.size(); ++j) v[j] = read_scalar(buf.data() + off[j], e.props[j].type);
        }
        if (!f) throw std::runtime_error(path + ": truncated PLY vertex data");
        positions[i] = {static_cast<float>(v[ix[0]]), static_cast<float>(v[ix[1]]), static_cast<float>(v[ix[2]])};
        if (hasN) normals[i] = {static_cast<float>(v[ix[3]]), static_cast<float>(v[ix[4]]), static_cast<float>(v[ix[5]])};
      }
      return;
    }
    throw std::runtime_error(path + ": PLY file without a vertex element");
  }
};

// scene::loadParameters (scene.cpp:1183-1250), field by field in the file's order.
struct ModelParams {
  float cameraQuat[4], cameraTranslation[3], cameraDist;
  float sigma, fixedSigmas[5], beginningSigma, factor, constant_shift, exponent, cutoff;
  uint32_t nbOfComputedSigmas, maxNbIterations, maxDepth, minDepth, method;
  int32_t kernelType;
  bool computedSigma, cut;
  uint32_t nbOfPoints, maxUpsamples;
  float offset[6];

  static constexpr size_t kBytes = 134;

  void load(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    char b[kBytes];
    if (!f || !f.read(b, kBytes)) throw std::runtime_error(path + ": not a 134-byte .params file");
    const char* p = b;
    auto get = [&p](void* dst, size_t n) {
      std::memcpy(dst, p, n);
      p += n;
    };
    get(cameraQuat, 16), get(cameraTranslation, 12), get(&cameraDist, 4);
    get(&sigma, 4), get(fixedSigmas, 20), get(&beginningSigma, 4), get(&factor, 4), get(&constant_shift, 4);
    get(&exponent, 4), get(&cutoff, 4);
    get(&nbOfComputedSigmas, 4), get(&maxNbIterations, 4), get(&maxDepth, 4), get(&minDepth, 4), get(&method, 4);
    get(&kernelType, 4);
    uint8_t c0, c1;
    get(&c0, 1), get(&c1, 1);
    computedSigma = c0 != 0, cut = c1 != 0;
    get(&nbOfPoints, 4), get(&maxUpsamples, 4), get(offset, 24);
    if (kernelType < 0 || kernelType > 2) throw std::runtime_error(path + ": unknown kernel type");
  }

#ifdef MLOD_APSS_HPP
  // The kernel of these settings for a cloud of bounding-box radius `radius`.
  Kernel kernel(float radius) const {
    Kernel k;
    k.type = static_cast<Kernel::KernelType>(kernelType);
    k.sigma = sigma;
    k.exponent = exponent;
    k.constant_shift = constant_shift;
    k.beginningSigma = beginningSigma;
    k.factor = factor;
    k.nbOfComputedSigmas = nbOfComputedSigmas;
    k.fixedSigmas.assign(fixedSigmas, fixedSigmas + 5);
    k.updateSigmas(computedSigma, radius);
    return k;
  }
#endif
};

}  // namespace mlod

#endif  // MLOD_PN_HPP

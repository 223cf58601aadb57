// LagrangianMesh.h -- a triangle cloth for the MPM substep (host side, no Eigen).
// Mirrors AnisotropicElastoplasticity/LagrangianMesh.h:13-117 for what the solver reads: vertex
// positions and velocities, faces, the fixed-vertex constraints (bindConstraints, :101) and the
// material (mu/lambda from Young's modulus and Poisson ratio, shear stiffness gamma, normal
// stiffness k, friction coefficient tan(phi)); and LagrangianMesh::ObjMesh (LagrangianMesh.cpp:
// 198-352), which reads an OBJ file. State is FP32 and planar: vertexPositions[3nv] = {x0[nv],
// x1[nv], x2[nv]}; faces are nf x 3 zero-based 32-bit vertex indices. Element masses, volumes and
// rest directions are derived on the device from the positions, density and thickness exactly as
// ObjMesh derives them (0.25 x area x thickness per element, LagrangianMesh.cpp:306-326).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mpm_c_abi.h"

namespace mpm {

class LagrangianMesh {
 public:
  std::vector<float> vertexPositions;   // [3nv] planar
  std::vector<float> vertexVelocities;  // [3nv] planar (zero after ObjMesh, LagrangianMesh.cpp:282-284)
  std::vector<int32_t> faces;           // [3nf] row-major, zero-based
  std::vector<uint8_t> vertexIsFixed;   // [nv] or empty (no constraints)
  std::vector<float> elementDirections;  // [9nf] D = [d1 d2 d3] row-major per element, filled by syncToHost

  double density = 2e3;
  double thickness = 0.04;
  double youngsModulus = 200.0;
  double poissonRatio = 0.3;
  double shearStiffness = 0.0;  // gamma
  double stiffness = 4e4;       // k
  double frictionAngleInDegree = 0.0;
  // 1: normal-force direction vector (D_rest^T).col(2), the reference's HybridSolver.cpp:432;
  // 0: D_rest^-T e3 (the two agree for a mesh in a z = const plane).
  bool referenceNormalRow = false;

  LagrangianMesh() = default;
  LagrangianMesh(std::vector<float> x, std::vector<int32_t> f, double density_, double thickness_, double E,
                 double nu, double gamma, double k, double phiDeg)
      : vertexPositions(std::move(x)), faces(std::move(f)), density(density_), thickness(thickness_),
        youngsModulus(E), poissonRatio(nu), shearStiffness(gamma), stiffness(k), frictionAngleInDegree(phiDeg) {
    if (vertexPositions.size() % 3 != 0) throw std::runtime_error("LagrangianMesh: positions must be 3 x nv");
    if (faces.size() % 3 != 0) throw std::runtime_error("LagrangianMesh: faces must be nf x 3");
    vertexVelocities.assign(vertexPositions.size(), 0.0f);
    validate();
  }

  size_t vertexCount() const { return vertexPositions.size() / 3; }
  size_t faceCount() const { return faces.size() / 3; }
  // Lame parameters and friction coefficient as ObjMesh computes them (LagrangianMesh.cpp:331-351).
  double mu() const { return youngsModulus / 2.0 / (1.0 + poissonRatio); }
  double lambda() const { return youngsModulus * poissonRatio / (1.0 + poissonRatio) / (1.0 - 2.0 * poissonRatio); }
  double frictionCoeff() const { return std::tan(frictionAngleInDegree * M_PI / 180.0); }

  // LagrangianMesh::bindConstraints (LagrangianMesh.cpp:355-365): one entry per vertex, nonzero =
  // fixed. A size mismatch raises instead of the reference's cerr + while(1).
  void bindConstraints(const std::vector<double>& vertexIsFixed_) {
    if (vertexIsFixed_.size() != vertexCount())
      throw std::runtime_error("LagrangianMesh::bindConstraints: the constraints are not compatible");
    vertexIsFixed.resize(vertexIsFixed_.size());
    for (size_t i = 0; i < vertexIsFixed_.size(); ++i) vertexIsFixed[i] = vertexIsFixed_[i] != 0.0 ? 1 : 0;
  }
  bool isFixed(size_t v) const { return !vertexIsFixed.empty() && vertexIsFixed[v] != 0; }

  mpm_mesh_params params() const {
    mpm_mesh_params p{};
    p.youngs_modulus = float(youngsModulus);
    p.poisson_ratio = float(poissonRatio);
    p.shear_stiffness = float(shearStiffness);
    p.stiffness = float(stiffness);
    p.friction_angle = float(frictionAngleInDegree);
    p.density = float(density);
    p.thickness = float(thickness);
    p.reference_normal_row = referenceNormalRow ? 1 : 0;
    return p;
  }

  // LagrangianMesh::ObjMesh (LagrangianMesh.cpp:198-352): "v x y z" and "f a b c" lines (1-based
  // indices; "a/b/c" and "a//c" forms keep the vertex index, negative indices count from the end
  // as in the OBJ format); every other line is skipped, as in the reference. A face with more than
  // three corners is split into a fan (the reference reads the first three only).
  static LagrangianMesh ObjMesh(const std::string& filename, double density, double thickness, double youngsModulus,
                                double poissonRatio, double shearStiffness, double stiffness,
                                double frictionAngleInDegree) {
    std::ifstream in(filename);
    if (!in) throw std::runtime_error("LagrangianMesh::ObjMesh: cannot open " + filename);
    std::vector<float> pos;  // interleaved while reading
    std::vector<int32_t> f;
    std::string line;
    size_t lineno = 0;
    while (std::getline(in, line)) {
      ++lineno;
      std::istringstream ls(line);
      std::string mark;
      if (!(ls >> mark)) continue;
      if (mark == "v") {
        double x, y, z;
        if (!(ls >> x >> y >> z)) throw std::runtime_error(filename + ":" + std::to_string(lineno) + ": bad vertex");
        pos.push_back(float(x));
        pos.push_back(float(y));
        pos.push_back(float(z));
      } else if (mark == "f") {
        std::vector<int64_t> corner;
        std::string tok;
        while (ls >> tok) {
          char* end = nullptr;
          const long long idx = std::strtoll(tok.c_str(), &end, 10);
          if (end == tok.c_str() || idx == 0)
            throw std::runtime_error(filename + ":" + std::to_string(lineno) + ": bad face index '" + tok + "'");
          const int64_t nvNow = int64_t(pos.size() / 3);
          const int64_t c = idx > 0 ? idx - 1 : nvNow + idx;
          if (c < 0 || c >= INT32_MAX)
            throw std::runtime_error(filename + ":" + std::to_string(lineno) + ": face index out of range '" + tok + "'");
          corner.push_back(c);
        }
        if (corner.size() < 3) throw std::runtime_error(filename + ":" + std::to_string(lineno) + ": face with < 3 corners");
        for (size_t c = 1; c + 1 < corner.size(); ++c) {
          f.push_back(int32_t(corner[0]));
          f.push_back(int32_t(corner[c]));
          f.push_back(int32_t(corner[c + 1]));
        }
      }
    }
    const size_t nv = pos.size() / 3;
    std::vector<float> planar(3 * nv);
    for (size_t i = 0; i < nv; ++i)
      for (int d = 0; d < 3; ++d) planar[d * nv + i] = pos[3 * i + d];
    return LagrangianMesh(std::move(planar), std::move(f), density, thickness, youngsModulus, poissonRatio,
                          shearStiffness, stiffness, frictionAngleInDegree);
  }

  // A flat nx x nz vertex sheet in the plane y = const with spacing `edge` from (x0, z0), two
  // triangles per quad, counter-clockwise seen from +y (normals up).
  static LagrangianMesh Grid(int nx, int nz, double edge, double x0, double z0, double y, double density,
                             double thickness, double youngsModulus, double poissonRatio, double shearStiffness,
                             double stiffness, double frictionAngleInDegree) {
    if (nx < 2 || nz < 2) throw std::runtime_error("LagrangianMesh::Grid needs at least 2 x 2 vertices");
    const size_t nv = size_t(nx) * size_t(nz);
    std::vector<float> x(3 * nv);
    for (int i = 0; i < nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const size_t v = size_t(i) * nz + k;
        x[v] = float(x0 + edge * i);
        x[nv + v] = float(y);
        x[2 * nv + v] = float(z0 + edge * k);
      }
    std::vector<int32_t> f;
    f.reserve(size_t(6) * (nx - 1) * (nz - 1));
    for (int i = 0; i + 1 < nx; ++i)
      for (int k = 0; k + 1 < nz; ++k) {
        const int32_t a = i * nz + k, b = (i + 1) * nz + k, d = i * nz + k + 1;
        f.insert(f.end(), {a, d, b});
      }
    for (int i = 0; i + 1 < nx; ++i)
      for (int k = 0; k + 1 < nz; ++k) {
        const int32_t b = (i + 1) * nz + k, c = (i + 1) * nz + k + 1, d = i * nz + k + 1;
        f.insert(f.end(), {b, d, c});
      }
    return LagrangianMesh(std::move(x), std::move(f), density, thickness, youngsModulus, poissonRatio, shearStiffness,
                          stiffness, frictionAngleInDegree);
  }

  // The reference's mesh frame file (HybridSolver.cpp:1006-1030): "v x y z" per vertex, then
  // "f a b c" per face with 1-based indices.
  void writeObj(const std::string& path) const {
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot write " + path);
    const size_t nv = vertexCount(), nf = faceCount();
    for (size_t v = 0; v < nv; ++v)
      out << "v " << vertexPositions[v] << ' ' << vertexPositions[nv + v] << ' ' << vertexPositions[2 * nv + v] << '\n';
    for (size_t e = 0; e < nf; ++e)
      out << "f " << faces[3 * e] + 1 << ' ' << faces[3 * e + 1] + 1 << ' ' << faces[3 * e + 2] + 1 << '\n';
    if (!out) throw std::runtime_error("write failed: " + path);
  }

  void validate() const {
    const int64_t nv = int64_t(vertexCount());
    if (nv > INT32_MAX) throw std::runtime_error("LagrangianMesh: more than 2^31 - 1 vertices");
    for (int32_t i : faces)
      if (i < 0 || i >= nv) throw std::runtime_error("LagrangianMesh: face index out of range");
    if (!vertexIsFixed.empty() && vertexIsFixed.size() != size_t(nv))
      throw std::runtime_error("LagrangianMesh: constraints size differs from the vertex count");
  }
};

}  // namespace mpm

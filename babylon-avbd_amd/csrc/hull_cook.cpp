// Convex hull cooking on the host: the hull a body collides as, from a point set or a closed triangulation, with its
// mass properties and principal frame, packed into the hull-buffer record the narrowphase reads.
//
// What is cooked follows reference/three-avbd/src/avbd3d/hull.ts (convexHull :60-66, hullFromTriangles :208-212):
//   * the convex hull of the distinct points (points equal to 1e-6 of the largest coordinate are one point); none when
//     they span no volume;
//   * coplanar neighbouring triangles merged into polygon faces (a face grows from a seed triangle over neighbours
//     within 1e-5 of the seed's plane and normal), each face one boundary loop, the surface closed (every edge on two
//     faces in opposite directions) with V - E + F = 2;
//   * at most 32 vertices: a bigger hull is replaced by the hull of 32 of its vertices picked by farthest-point
//     sampling from the vertex farthest from the centre of mass;
//   * mass properties per unit density from signed tetrahedra over the fan-triangulated faces; the principal frame
//     from the inertia tensor's eigenvectors (cyclic Jacobi, right-handed); vertices, face planes (offset = the largest
//     n·v over the face's vertices) and edges (a < b, the face on each side) in that frame, centred on the centre of
//     mass; full widths of the vertices' bounds along the principal axes and the farthest vertex's distance.
// The construction is an incremental hull: a tetrahedron of extreme points, then the other points farthest first,
// each removing the triangles it sees (grown from those it is clearly in front of across neighbours it is level with,
// so coplanar triangles go together) and closing the hole with a fan from its horizon.
//
// Record layout (wgsl-hull.ts:13-19; offsets in 16-byte words from the start of the buffer it is placed in):
//   h:   vertex start, vertex count, face start, face count
//   h+1: edge start, edge count, index start, 0
//   vertices (x, y, z, 0 as f32 bits), faces (normal.xyz + offset as f32 bits; first index, index count, 0, 0),
//   vertex indices 4 per word, edges (a, b, face with a→b, face with b→a).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "hull_cook.h"

namespace avbd {
namespace hullcook {
namespace {

using V = std::array<double, 3>;

inline V vsub(const V &a, const V &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline double vdot(const V &a, const V &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V vcross(const V &a, const V &b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double vlen(const V &a) { return std::sqrt(vdot(a, a)); }
inline uint64_t edgeKey(uint32_t a, uint32_t b) { return uint64_t(a) << 32 | b; }

struct Tri {
  uint32_t v[3];
  V n;
  double d;
  bool alive = true;
};

Tri makeTri(const std::vector<V> &p, uint32_t a, uint32_t b, uint32_t c) {
  V n = vcross(vsub(p[b], p[a]), vsub(p[c], p[a]));
  double l = vlen(n);
  if (!(l > 0)) l = 1;
  Tri t;
  t.v[0] = a;
  t.v[1] = b;
  t.v[2] = c;
  t.n = {n[0] / l, n[1] / l, n[2] / l};
  t.d = vdot(t.n, p[a]);
  return t;
}

/** Eigen-decomposition of a symmetric 3x3 matrix by cyclic Jacobi rotations; columns of `vec` are the eigenvectors. */
void jacobiEigen(double a[3][3], double val[3], double vec[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) vec[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; sweep++) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off < 1e-30 * (diag + 1e-300)) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        if (std::fabs(a[p][q]) < 1e-300) continue;
        // The rotation that zeroes a[p][q]: tan of its angle is the smaller root of t² + 2θt - 1 = 0
        const double theta = (a[q][q] - a[p][p]) / (2 * a[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < 3; k++) {  // columns p, q
          const double kp = a[k][p], kq = a[k][q];
          a[k][p] = c * kp - s * kq;
          a[k][q] = s * kp + c * kq;
        }
        for (int k = 0; k < 3; k++) {  // rows p, q
          const double pk = a[p][k], qk = a[q][k];
          a[p][k] = c * pk - s * qk;
          a[q][k] = s * pk + c * qk;
        }
        for (int k = 0; k < 3; k++) {
          const double kp = vec[k][p], kq = vec[k][q];
          vec[k][p] = c * kp - s * kq;
          vec[k][q] = s * kp + c * kq;
        }
      }
  }
  for (int i = 0; i < 3; i++) val[i] = a[i][i];
  // Right-handed: the third axis is the first × the second
  const V x = {vec[0][0], vec[1][0], vec[2][0]}, y = {vec[0][1], vec[1][1], vec[2][1]};
  const V z = vcross(x, y);
  for (int k = 0; k < 3; k++) vec[k][2] = z[k];
}

/** Unit quaternion (x, y, z, w) of the rotation whose matrix (rows) is m (Shepperd's method). */
void quatOfMatrix(const double m[3][3], double q[4]) {
  const double tr = m[0][0] + m[1][1] + m[2][2];
  double x, y, z, w;
  if (tr > 0) {
    const double s = 2 * std::sqrt(tr + 1);
    w = s / 4;
    x = (m[2][1] - m[1][2]) / s;
    y = (m[0][2] - m[2][0]) / s;
    z = (m[1][0] - m[0][1]) / s;
  } else if (m[0][0] > m[1][1] && m[0][0] > m[2][2]) {
    const double s = 2 * std::sqrt(1 + m[0][0] - m[1][1] - m[2][2]);
    x = s / 4;
    w = (m[2][1] - m[1][2]) / s;
    y = (m[0][1] + m[1][0]) / s;
    z = (m[0][2] + m[2][0]) / s;
  } else if (m[1][1] > m[2][2]) {
    const double s = 2 * std::sqrt(1 + m[1][1] - m[0][0] - m[2][2]);
    y = s / 4;
    w = (m[0][2] - m[2][0]) / s;
    x = (m[0][1] + m[1][0]) / s;
    z = (m[1][2] + m[2][1]) / s;
  } else {
    const double s = 2 * std::sqrt(1 + m[2][2] - m[0][0] - m[1][1]);
    z = s / 4;
    w = (m[1][0] - m[0][1]) / s;
    x = (m[0][2] + m[2][0]) / s;
    y = (m[1][2] + m[2][1]) / s;
  }
  const double l = std::sqrt(x * x + y * y + z * z + w * w);
  q[0] = x / l;
  q[1] = y / l;
  q[2] = z / l;
  q[3] = w / l;
}

/** Faces, closure checks, mass properties and the principal frame of a closed outward triangulation of points p. */
bool fromTriangulation(const std::vector<V> &p, const std::vector<std::array<uint32_t, 3>> &triangles, Shape &out,
                       std::string &err) {
  double scale = 0;
  for (const V &v : p) scale = std::max({scale, std::fabs(v[0]), std::fabs(v[1]), std::fabs(v[2])});
  const double planeTol = std::max(scale, 1e-9) * 1e-5;
  std::vector<Tri> tris;
  for (const auto &t : triangles) {
    Tri k = makeTri(p, t[0], t[1], t[2]);
    const V c = vcross(vsub(p[t[1]], p[t[0]]), vsub(p[t[2]], p[t[0]]));
    if (!(vlen(c) > 0)) continue;  // degenerate triangles carry no surface
    tris.push_back(k);
  }
  std::unordered_map<uint64_t, uint32_t> triOfEdge;  // directed edge -> triangle
  for (uint32_t k = 0; k < tris.size(); k++)
    for (int j = 0; j < 3; j++) triOfEdge[edgeKey(tris[k].v[j], tris[k].v[(j + 1) % 3])] = k;

  // Faces: regions grown from a seed triangle over neighbours lying on the seed's plane
  std::vector<int> groupOf(tris.size(), -1);
  std::vector<std::vector<uint32_t>> groups;
  for (uint32_t k = 0; k < tris.size(); k++) {
    if (groupOf[k] >= 0) continue;
    const Tri &seed = tris[k];
    std::vector<uint32_t> members{k};
    groupOf[k] = int(groups.size());
    for (size_t m = 0; m < members.size(); m++) {
      const Tri &t = tris[members[m]];
      for (int j = 0; j < 3; j++) {
        auto it = triOfEdge.find(edgeKey(t.v[(j + 1) % 3], t.v[j]));
        if (it == triOfEdge.end() || groupOf[it->second] >= 0) continue;
        const Tri &u = tris[it->second];
        bool onPlane = vdot(seed.n, u.n) > 1 - 1e-5;
        for (int s = 0; s < 3 && onPlane; s++) onPlane = std::fabs(vdot(seed.n, p[u.v[s]]) - seed.d) <= planeTol;
        if (!onPlane) continue;
        groupOf[it->second] = int(groups.size());
        members.push_back(it->second);
      }
    }
    groups.push_back(std::move(members));
  }

  struct Poly {
    V n;
    std::vector<uint32_t> verts;
  };
  std::vector<Poly> polys;
  for (const auto &members : groups) {
    std::unordered_set<uint64_t> inGroup;
    for (uint32_t k : members)
      for (int j = 0; j < 3; j++) inGroup.insert(edgeKey(tris[k].v[j], tris[k].v[(j + 1) % 3]));
    // Boundary edges (those whose reverse is not in the face), as a successor map in first-seen order
    std::vector<std::pair<uint32_t, uint32_t>> boundary;
    std::unordered_map<uint32_t, uint32_t> next;
    for (uint32_t k : members)
      for (int j = 0; j < 3; j++) {
        const uint32_t a = tris[k].v[j], b = tris[k].v[(j + 1) % 3];
        if (inGroup.count(edgeKey(b, a))) continue;
        if (next.count(a)) {
          err = "a face's boundary is not one loop";
          return false;
        }
        next[a] = b;
        boundary.emplace_back(a, b);
      }
    V n = {0, 0, 0};  // area-weighted normal
    for (uint32_t k : members) {
      const Tri &t = tris[k];
      const V c = vcross(vsub(p[t.v[1]], p[t.v[0]]), vsub(p[t.v[2]], p[t.v[0]]));
      for (int s = 0; s < 3; s++) n[s] += c[s];
    }
    double l = vlen(n);
    if (!(l > 0)) l = 1;
    Poly poly;
    poly.n = {n[0] / l, n[1] / l, n[2] / l};
    const uint32_t start = boundary.front().first;
    poly.verts.push_back(start);
    for (uint32_t v = next[start]; v != start && poly.verts.size() <= next.size(); v = next[v]) {
      if (!next.count(v)) {
        err = "a face's boundary is not closed";
        return false;
      }
      poly.verts.push_back(v);
    }
    if (poly.verts.size() != next.size()) {
      err = "a face's boundary is not one loop";
      return false;
    }
    polys.push_back(std::move(poly));
  }

  // The vertices the faces use, numbered in order of first use
  std::unordered_map<uint32_t, uint32_t> index;
  std::vector<V> verts;
  for (Poly &poly : polys)
    for (uint32_t &v : poly.verts) {
      auto it = index.find(v);
      if (it == index.end()) {
        it = index.emplace(v, uint32_t(verts.size())).first;
        verts.push_back(p[v]);
      }
      v = it->second;
    }

  // Closed: every directed edge once, its reverse present; V - E + F = 2
  std::unordered_set<uint64_t> directed;
  size_t sides = 0;
  for (const Poly &poly : polys)
    for (size_t k = 0; k < poly.verts.size(); k++) {
      directed.insert(edgeKey(poly.verts[k], poly.verts[(k + 1) % poly.verts.size()]));
      sides++;
    }
  bool closed = directed.size() == sides;
  for (uint64_t e : directed)
    if (closed && !directed.count(edgeKey(uint32_t(e & 0xffffffffu), uint32_t(e >> 32)))) closed = false;
  if (!closed || int64_t(verts.size()) - int64_t(directed.size() / 2) + int64_t(polys.size()) != 2) {
    err = "the triangles do not close up";
    return false;
  }

  // Mass properties per unit density: signed tetrahedra (origin, a, b, c) over each face's fan
  double volume = 0;
  V com = {0, 0, 0};
  double C[6] = {0, 0, 0, 0, 0, 0};  // ∫ xx, yy, zz, xy, xz, yz
  for (const Poly &poly : polys) {
    const V &a = verts[poly.verts[0]];
    for (size_t k = 1; k + 1 < poly.verts.size(); k++) {
      const V &b = verts[poly.verts[k]], &c = verts[poly.verts[k + 1]];
      const double det = vdot(a, vcross(b, c));
      volume += det / 6;
      for (int j = 0; j < 3; j++) com[j] += det / 24 * (a[j] + b[j] + c[j]);
      // ∫ x_i x_j over the tetrahedron = det/120 (Σ_v v_i v_j + (Σ_v v_i)(Σ_v v_j)), the origin vertex being zero
      auto second = [&](int i, int j) {
        return det / 120 * (a[i] * a[j] + b[i] * b[j] + c[i] * c[j] + (a[i] + b[i] + c[i]) * (a[j] + b[j] + c[j]));
      };
      C[0] += second(0, 0);
      C[1] += second(1, 1);
      C[2] += second(2, 2);
      C[3] += second(0, 1);
      C[4] += second(0, 2);
      C[5] += second(1, 2);
    }
  }
  if (!(volume > 0)) {
    err = "the surface encloses no volume";
    return false;
  }
  for (int j = 0; j < 3; j++) com[j] /= volume;
  C[0] -= volume * com[0] * com[0];
  C[1] -= volume * com[1] * com[1];
  C[2] -= volume * com[2] * com[2];
  C[3] -= volume * com[0] * com[1];
  C[4] -= volume * com[0] * com[2];
  C[5] -= volume * com[1] * com[2];
  const double trace = C[0] + C[1] + C[2];
  double I[3][3] = {{trace - C[0], -C[3], -C[4]}, {-C[3], trace - C[1], -C[5]}, {-C[4], -C[5], trace - C[2]}};
  double moments[3], R[3][3];
  jacobiEigen(I, moments, R);

  // Principal frame: local = Rᵀ (x - com)
  auto toLocal = [&](const V &x) {
    const V d = vsub(x, com);
    return V{R[0][0] * d[0] + R[1][0] * d[1] + R[2][0] * d[2], R[0][1] * d[0] + R[1][1] * d[1] + R[2][1] * d[2],
             R[0][2] * d[0] + R[1][2] * d[1] + R[2][2] * d[2]};
  };
  std::vector<V> local;
  for (const V &v : verts) local.push_back(toLocal(v));
  out = Shape();
  out.vertices.reserve(local.size() * 3);
  for (const V &v : local)
    for (int k = 0; k < 3; k++) out.vertices.push_back(float(v[k]));
  for (const Poly &poly : polys) {
    Face f;
    const V &n = poly.n;
    f.normal = {R[0][0] * n[0] + R[1][0] * n[1] + R[2][0] * n[2], R[0][1] * n[0] + R[1][1] * n[1] + R[2][1] * n[2],
                R[0][2] * n[0] + R[1][2] * n[1] + R[2][2] * n[2]};
    f.d = -std::numeric_limits<double>::infinity();
    for (uint32_t v : poly.verts) f.d = std::max(f.d, vdot(f.normal, local[v]));
    f.verts = poly.verts;
    out.faces.push_back(std::move(f));
  }
  std::unordered_map<uint64_t, uint32_t> faceOf;
  for (uint32_t k = 0; k < out.faces.size(); k++) {
    const auto &fv = out.faces[k].verts;
    for (size_t j = 0; j < fv.size(); j++) faceOf[edgeKey(fv[j], fv[(j + 1) % fv.size()])] = k;
  }
  for (uint32_t k = 0; k < out.faces.size(); k++) {
    const auto &fv = out.faces[k].verts;
    for (size_t j = 0; j < fv.size(); j++) {
      const uint32_t a = fv[j], b = fv[(j + 1) % fv.size()];
      auto it = faceOf.find(edgeKey(b, a));
      if (a < b && it != faceOf.end()) out.edges.push_back({a, b, k, it->second});
    }
  }
  out.volume = volume;
  for (int k = 0; k < 3; k++) out.moments[k] = moments[k];
  out.size[0] = out.size[1] = out.size[2] = 0;
  out.radius = 0;
  for (const V &v : local) {
    for (int k = 0; k < 3; k++) out.size[k] = std::max(out.size[k], 2 * std::fabs(v[k]));
    out.radius = std::max(out.radius, vlen(v));
  }
  for (int k = 0; k < 3; k++) out.center[k] = com[k];
  quatOfMatrix(R, out.rotation);
  return true;
}

/** The convex hull of xyz points (no simplification). */
bool buildHull(const std::vector<V> &input, Shape &out, std::string &err) {
  double scale = 0;
  for (const V &v : input) scale = std::max({scale, std::fabs(v[0]), std::fabs(v[1]), std::fabs(v[2])});
  const double eps = std::max(scale, 1e-9) * 1e-6;
  // Distinct points (split vertices repeat positions): one per cell of size eps
  struct KeyHash {
    size_t operator()(const std::array<int64_t, 3> &k) const {
      return size_t(uint64_t(k[0]) * 0x9E3779B97F4A7C15ull ^ uint64_t(k[1]) * 0xC2B2AE3D27D4EB4Full ^
                    uint64_t(k[2]) * 0x165667B19E3779F9ull);
    }
  };
  std::unordered_set<std::array<int64_t, 3>, KeyHash> seen;
  std::vector<V> p;
  for (const V &v : input) {
    const std::array<int64_t, 3> key = {int64_t(std::llround(v[0] / eps)), int64_t(std::llround(v[1] / eps)),
                                        int64_t(std::llround(v[2] / eps))};
    if (seen.insert(key).second) p.push_back(v);
  }
  err = "the points span no volume";
  if (p.size() < 4) return false;
  const double tol = std::max(scale, 1e-9) * 1e-10;

  // A tetrahedron of extreme points: lowest x; farthest from it; farthest from their line; farthest from their plane
  uint32_t i0 = 0;
  for (uint32_t i = 1; i < p.size(); i++)
    if (p[i][0] < p[i0][0]) i0 = i;
  auto argmax = [&](auto score) {
    int64_t at = -1;
    double best = 0;
    for (uint32_t i = 0; i < p.size(); i++) {
      const double s = score(i);
      if (s > best) {
        best = s;
        at = i;
      }
    }
    return std::make_pair(at, best);
  };
  auto [i1, d1] = argmax([&](uint32_t i) { return vlen(vsub(p[i], p[i0])); });
  if (i1 < 0 || d1 < tol) return false;
  const V e01 = vsub(p[size_t(i1)], p[i0]);
  auto [i2, d2] = argmax([&](uint32_t i) { return vlen(vcross(e01, vsub(p[i], p[i0]))); });
  if (i2 < 0 || d2 < tol * vlen(e01)) return false;
  const V n012 = vcross(e01, vsub(p[size_t(i2)], p[i0]));
  auto [i3, d3] = argmax([&](uint32_t i) { return std::fabs(vdot(n012, vsub(p[i], p[i0]))); });
  if (i3 < 0 || d3 < tol * vlen(n012)) return false;
  const uint32_t a = i0, b = uint32_t(i1), c = uint32_t(i2), d = uint32_t(i3);

  std::vector<Tri> tris;
  const bool flip = vdot(n012, vsub(p[d], p[a])) > 0;  // the base faces away from the fourth point
  const uint32_t b0 = a, b1 = flip ? c : b, b2 = flip ? b : c;
  tris.push_back(makeTri(p, b0, b1, b2));
  tris.push_back(makeTri(p, b0, d, b1));
  tris.push_back(makeTri(p, b1, d, b2));
  tris.push_back(makeTri(p, b2, d, b0));

  // The other points, farthest from the tetrahedron's centre first (stable for ties)
  const V centre = {(p[a][0] + p[b][0] + p[c][0] + p[d][0]) / 4, (p[a][1] + p[b][1] + p[c][1] + p[d][1]) / 4,
                    (p[a][2] + p[b][2] + p[c][2] + p[d][2]) / 4};
  std::vector<uint32_t> order;
  for (uint32_t i = 0; i < p.size(); i++)
    if (i != a && i != b && i != c && i != d) order.push_back(i);
  std::vector<double> dist(p.size());
  for (uint32_t i : order) dist[i] = vlen(vsub(p[i], centre));
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return dist[x] > dist[y]; });

  std::unordered_map<uint64_t, uint32_t> triOfEdge;  // directed edge -> live triangle
  auto link = [&](uint32_t k) {
    for (int j = 0; j < 3; j++) triOfEdge[edgeKey(tris[k].v[j], tris[k].v[(j + 1) % 3])] = k;
  };
  for (uint32_t k = 0; k < tris.size(); k++) link(k);
  std::vector<uint32_t> region, stack;
  std::vector<char> inRegion;
  for (uint32_t i : order) {
    region.clear();
    inRegion.assign(tris.size(), 0);
    for (uint32_t k = 0; k < tris.size(); k++)
      if (tris[k].alive && vdot(tris[k].n, p[i]) - tris[k].d > tol) {
        region.push_back(k);
        inRegion[k] = 1;
      }
    if (region.empty()) continue;  // inside (or on) the hull
    // Grow across neighbours the point is level with or in front of, so coplanar triangles leave together
    stack.assign(region.begin(), region.end());
    while (!stack.empty()) {
      const uint32_t k = stack.back();
      stack.pop_back();
      for (int j = 0; j < 3; j++) {
        auto it = triOfEdge.find(edgeKey(tris[k].v[(j + 1) % 3], tris[k].v[j]));
        if (it == triOfEdge.end()) continue;
        const uint32_t u = it->second;
        if (inRegion[u] || !tris[u].alive || !(vdot(tris[u].n, p[i]) - tris[u].d > -tol)) continue;
        inRegion[u] = 1;
        region.push_back(u);
        stack.push_back(u);
      }
    }
    // Horizon: directed edges of the removed triangles whose reverse belongs to a kept one
    std::vector<std::pair<uint32_t, uint32_t>> horizon;
    for (uint32_t k : region)
      for (int j = 0; j < 3; j++) {
        const uint32_t ea = tris[k].v[j], eb = tris[k].v[(j + 1) % 3];
        auto it = triOfEdge.find(edgeKey(eb, ea));
        if (it == triOfEdge.end() || !inRegion[it->second]) horizon.emplace_back(ea, eb);
      }
    for (uint32_t k : region) {
      tris[k].alive = false;
      for (int j = 0; j < 3; j++) {
        auto it = triOfEdge.find(edgeKey(tris[k].v[j], tris[k].v[(j + 1) % 3]));
        if (it != triOfEdge.end() && it->second == k) triOfEdge.erase(it);
      }
    }
    for (const auto &[ea, eb] : horizon) {
      tris.push_back(makeTri(p, ea, eb, i));
      link(uint32_t(tris.size() - 1));
    }
  }
  std::vector<std::array<uint32_t, 3>> live;
  for (const Tri &t : tris)
    if (t.alive) live.push_back({t.v[0], t.v[1], t.v[2]});
  return fromTriangulation(p, live, out, err);
}

/** `count` of a hull's vertices, in the input points' frame, picked by farthest-point sampling. */
std::vector<V> spreadVertices(const Shape &h, size_t count) {
  const double x = h.rotation[0], y = h.rotation[1], z = h.rotation[2], w = h.rotation[3];
  std::vector<V> verts;
  for (size_t i = 0; i + 2 < h.vertices.size(); i += 3) {
    const V v = {h.vertices[i], h.vertices[i + 1], h.vertices[i + 2]};
    // v + w t + q × t with t = 2 q × v: the principal frame's rotation applied, then the centre added back
    const V t = {2 * (y * v[2] - z * v[1]), 2 * (z * v[0] - x * v[2]), 2 * (x * v[1] - y * v[0])};
    verts.push_back({v[0] + w * t[0] + (y * t[2] - z * t[1]) + h.center[0],
                     v[1] + w * t[1] + (z * t[0] - x * t[2]) + h.center[1],
                     v[2] + w * t[2] + (x * t[1] - y * t[0]) + h.center[2]});
  }
  const V c = {h.center[0], h.center[1], h.center[2]};
  size_t first = 0;
  for (size_t i = 0; i < verts.size(); i++)
    if (vlen(vsub(verts[i], c)) > vlen(vsub(verts[first], c))) first = i;
  std::vector<size_t> chosen{first};
  std::vector<double> dist(verts.size());
  for (size_t i = 0; i < verts.size(); i++) dist[i] = vlen(vsub(verts[i], verts[first]));
  while (chosen.size() < count) {
    size_t far = 0;
    for (size_t i = 0; i < verts.size(); i++)
      if (dist[i] > dist[far]) far = i;
    chosen.push_back(far);
    for (size_t i = 0; i < verts.size(); i++) dist[i] = std::min(dist[i], vlen(vsub(verts[i], verts[far])));
  }
  std::vector<V> out;
  for (size_t i : chosen) out.push_back(verts[i]);
  return out;
}

std::vector<V> toPoints(const double *xyz, size_t n) {
  std::vector<V> p(n);
  for (size_t i = 0; i < n; i++) p[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  return p;
}

}  // namespace

bool fromPoints(const double *xyz, size_t n, Shape &out, std::string &err) {
  Shape full;
  if (!buildHull(toPoints(xyz, n), full, err)) return false;
  if (full.vertices.size() / 3 <= MAX_HULL_VERTICES) {
    out = std::move(full);
    return true;
  }
  return buildHull(spreadVertices(full, MAX_HULL_VERTICES), out, err);
}

bool fromTriangles(const double *xyz, size_t n, const uint32_t *indices, size_t nIndices, Shape &out,
                   std::string &err) {
  if (nIndices % 3) {
    err = "indices must come in triangles";
    return false;
  }
  std::vector<std::array<uint32_t, 3>> tris;
  for (size_t i = 0; i < nIndices; i += 3) {
    for (int k = 0; k < 3; k++)
      if (indices[i + k] >= n) {
        err = "a triangle index is out of range";
        return false;
      }
    tris.push_back({indices[i], indices[i + 1], indices[i + 2]});
  }
  Shape h;
  if (!fromTriangulation(toPoints(xyz, n), tris, h, err)) return false;
  if (h.vertices.size() / 3 <= MAX_HULL_VERTICES) {
    out = std::move(h);
    return true;
  }
  return buildHull(spreadVertices(h, MAX_HULL_VERTICES), out, err);
}

size_t recordWords(const Shape &s) {
  size_t indexCount = 0;
  for (const Face &f : s.faces) indexCount += f.verts.size();
  return 2 + s.vertices.size() / 3 + 2 * s.faces.size() + (indexCount + 3) / 4 + s.edges.size();
}

size_t pack(const Shape &s, uint32_t at, uint32_t *out) {
  const uint32_t V = uint32_t(s.vertices.size() / 3), F = uint32_t(s.faces.size()), E = uint32_t(s.edges.size());
  size_t indexCount = 0;
  for (const Face &f : s.faces) indexCount += f.verts.size();
  const uint32_t I = uint32_t((indexCount + 3) / 4);
  const uint32_t vs = at + 2, fs = vs + V, is = fs + 2 * F, es = is + I;
  const size_t words = recordWords(s);
  std::memset(out, 0, words * 16);
  auto fbits = [](float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return u;
  };
  uint32_t *h = out;
  h[0] = vs;
  h[1] = V;
  h[2] = fs;
  h[3] = F;
  h[4] = es;
  h[5] = E;
  h[6] = is;
  h[7] = 0;
  for (uint32_t v = 0; v < V; v++)
    for (int k = 0; k < 3; k++) out[(2 + v) * 4 + k] = fbits(s.vertices[3 * v + k]);
  uint32_t index = 0;
  for (uint32_t k = 0; k < F; k++) {
    const Face &f = s.faces[k];
    uint32_t *o = out + (2 + V + 2 * k) * 4;
    for (int c = 0; c < 3; c++) o[c] = fbits(float(f.normal[c]));
    o[3] = fbits(float(f.d));
    o[4] = index;
    o[5] = uint32_t(f.verts.size());
    for (uint32_t v : f.verts) out[(2 + V + 2 * F) * 4 + index++] = v;
  }
  for (uint32_t k = 0; k < E; k++)
    for (int c = 0; c < 4; c++) out[(2 + V + 2 * F + I + k) * 4 + c] = s.edges[k][c];
  return words;
}

}  // namespace hullcook
}  // namespace avbd

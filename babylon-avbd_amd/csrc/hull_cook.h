// Host-side convex hull cooking (hull_cook.cpp): shapes and the packed hull-buffer record.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace avbd {
namespace hullcook {

/** Most vertices a cooked hull keeps (reference/three-avbd/src/avbd3d/hull.ts:48-54: feature keys hold 8-bit indices). */
constexpr size_t MAX_HULL_VERTICES = 32;
/** Largest record of a hull of at most 32 vertices (2 + 32 + 2·60 + ⌈180/4⌉ + 90 words, rounded up). */
constexpr size_t MAX_RECORD_WORDS = 320;

struct Face {
  std::array<double, 3> normal;  // outward unit normal, principal frame
  double d;                      // plane offset: normal · x = d
  std::vector<uint32_t> verts;   // counter-clockwise seen from outside
};

struct Shape {
  std::vector<float> vertices;                 // x, y, z per vertex, principal frame centred on the centre of mass
  std::vector<Face> faces;
  std::vector<std::array<uint32_t, 4>> edges;  // a < b, face with a→b, face with b→a
  double volume = 0;
  double moments[3] = {0, 0, 0};   // principal moments per unit density
  double size[3] = {0, 0, 0};      // full widths of the vertices' bounds along the principal axes
  double radius = 0;               // farthest vertex from the centre of mass
  double center[3] = {0, 0, 0};    // the principal frame in the input frame: centre of mass ...
  double rotation[4] = {0, 0, 0, 1};  // ... and rotation (x, y, z, w)
};

bool fromPoints(const double *xyz, size_t n, Shape &out, std::string &err);
bool fromTriangles(const double *xyz, size_t n, const uint32_t *indices, size_t nIndices, Shape &out, std::string &err);
/** Size of the shape's record in 16-byte words. */
size_t recordWords(const Shape &s);
/** The record (4 u32 per word) with its header at word `at` of the buffer it goes in; returns its size in words. */
size_t pack(const Shape &s, uint32_t at, uint32_t *out);

}  // namespace hullcook
}  // namespace avbd

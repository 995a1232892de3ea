// Host-built acceleration structure replacing Embree's rtcCommitScene
// (core/src/Scene.cpp:191): a binned-SAH binary tree collapsed into a 4- or
// 8-wide BVH, with leaf triangles stored as 48-byte records in leaf order.
// A 4-wide node is 128 bytes; an 8-wide node is two such halves back to back
// (256 bytes, children 0-3 in the first half, 4-7 in the second), so that
// the traversal tests an 8-wide node as two 4-wide ones.
//
// Node (half) layout (8 x 16 bytes):
//   [0] lo.x of children 0..3   [1] hi.x   [2] lo.y   [3] hi.y
//   [4] lo.z                    [5] hi.z   [6] child refs (int32 x4)
//   [7] reserved
// child ref >= 0: index of an inner node
// child ref <  0: 0x80000000 | count << 24 | first leaf triangle (count 1..127);
//                 count 0 marks an empty slot (all six planes +inf: a slab
//                 test gives it entry distance +inf or exit distance -inf,
//                 so the traversal never enters it and needs no ref test)
// Triangle record (3 x 16 bytes): v0.xyz + global triangle index (bits),
// v1.xyz + 0, v2.xyz + 0.
//
// Boxes are widened by 2^-18 of their magnitude so that conservative slab
// tests never drop a triangle the watertight test would report.
#pragma once
#include <stdint.h>

#include <vector>

namespace hasty {

struct BvhBuildOptions {
  int binCount = 32;
  int maxLeafSize = 4;      // triangles per leaf
  float traversalCost = 1.0f;
  float intersectCost = 1.0f;
};

struct Bvh4 {
  int width = 4;                // children per node: 4 or 8
  std::vector<float> nodes;     // 8 * width floats (32 * width bytes) per node
  std::vector<float> tris;      // 12 floats (48 B) per leaf triangle
  uint32_t numNodes = 0;
  uint32_t numTris = 0;
  int depth = 0;                // inner levels on the longest root-to-leaf path
  int binaryDepth = 0;
  uint64_t binaryNodes = 0;
};

// triPos: 9 floats per triangle (global order); width 4 or 8.
Bvh4 buildBvh4(const float* triPos, uint32_t numTris, const BvhBuildOptions& opt, int width = 4);

constexpr uint32_t kLeafFlag = 0x80000000u;
inline uint32_t makeLeafRef(uint32_t first, uint32_t count) { return kLeafFlag | (count << 24) | first; }

}  // namespace hasty

// Binned-SAH build + 4-wide collapse; see bvh.h for the layout.
#include "bvh.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <future>
#include <stdexcept>
#include <thread>

namespace hasty {
namespace {

struct Box {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  void grow(const Box& b) {
    for (int c = 0; c < 3; ++c) {
      lo[c] = std::min(lo[c], b.lo[c]);
      hi[c] = std::max(hi[c], b.hi[c]);
    }
  }
  void growPoint(const float* p) {
    for (int c = 0; c < 3; ++c) {
      lo[c] = std::min(lo[c], p[c]);
      hi[c] = std::max(hi[c], p[c]);
    }
  }
  float area() const {
    if (lo[0] > hi[0]) return 0.0f;
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
  }
};

struct BNode {  // binary node
  Box box;
  int32_t child[2] = {-1, -1};
  uint32_t first = 0, count = 0;  // leaf range in `order`
};

struct Builder {
  const BvhBuildOptions& opt;
  std::vector<Box> triBox;
  std::vector<float> cen;  // 3 per triangle
  std::vector<uint32_t> order;
  std::vector<BNode> nodes;
  std::mutex mtx;

  explicit Builder(const BvhBuildOptions& o) : opt(o) {}

  int32_t alloc(int n) {
    std::lock_guard<std::mutex> g(mtx);
    int32_t i = (int32_t)nodes.size();
    for (int k = 0; k < n; ++k) nodes.emplace_back();
    return i;
  }

  // Builds the subtree over order[first, first+count) into node `ni`; returns depth.
  int build(int32_t ni, uint32_t first, uint32_t count, int depthNow) {
    Box box, cbox;
    for (uint32_t k = first; k < first + count; ++k) {
      uint32_t t = order[k];
      box.grow(triBox[t]);
      cbox.growPoint(&cen[(size_t)t * 3]);
    }
    auto makeLeaf = [&]() {
      std::lock_guard<std::mutex> g(mtx);
      nodes[ni].box = box;
      nodes[ni].first = first;
      nodes[ni].count = count;
      return 1;
    };
    if (count == 1) return makeLeaf();
    const int NB = opt.binCount;
    float bestCost = FLT_MAX;
    int bestAxis = -1, bestSplit = -1;
    std::vector<Box> bins(NB), rightBox(NB);
    std::vector<uint32_t> cnt(NB), rightCnt(NB);
    for (int axis = 0; axis < 3; ++axis) {
      float ext = cbox.hi[axis] - cbox.lo[axis];
      if (!(ext > 0.0f)) continue;
      float scale = NB / ext;
      std::fill(cnt.begin(), cnt.end(), 0u);
      for (int b = 0; b < NB; ++b) bins[b] = Box();
      for (uint32_t k = first; k < first + count; ++k) {
        uint32_t t = order[k];
        int b = (int)((cen[(size_t)t * 3 + axis] - cbox.lo[axis]) * scale);
        b = std::min(std::max(b, 0), NB - 1);
        cnt[b]++;
        bins[b].grow(triBox[t]);
      }
      Box acc;
      uint32_t ac = 0;
      for (int b = NB - 1; b > 0; --b) {
        acc.grow(bins[b]);
        ac += cnt[b];
        rightBox[b] = acc;
        rightCnt[b] = ac;
      }
      Box lacc;
      uint32_t lc = 0;
      for (int b = 0; b < NB - 1; ++b) {
        lacc.grow(bins[b]);
        lc += cnt[b];
        if (lc == 0 || rightCnt[b + 1] == 0) continue;
        float cost = lacc.area() * lc + rightBox[b + 1].area() * rightCnt[b + 1];
        if (cost < bestCost) {
          bestCost = cost;
          bestAxis = axis;
          bestSplit = b;
        }
      }
    }
    float parentArea = box.area();
    float splitCost = opt.traversalCost + opt.intersectCost * (parentArea > 0.0f ? bestCost / parentArea : (float)count);
    float leafCost = opt.intersectCost * count;
    if ((int)count <= opt.maxLeafSize && (bestAxis < 0 || leafCost <= splitCost)) return makeLeaf();
    uint32_t mid;
    if (bestAxis < 0) {
      mid = first + count / 2;  // coincident centroids
    } else {
      float ext = cbox.hi[bestAxis] - cbox.lo[bestAxis];
      float scale = NB / ext;
      auto it = std::partition(order.begin() + first, order.begin() + first + count, [&](uint32_t t) {
        int b = (int)((cen[(size_t)t * 3 + bestAxis] - cbox.lo[bestAxis]) * scale);
        b = std::min(std::max(b, 0), NB - 1);
        return b <= bestSplit;
      });
      mid = (uint32_t)(it - order.begin());
      if (mid == first || mid == first + count) mid = first + count / 2;
    }
    int32_t c0 = alloc(2);
    {
      std::lock_guard<std::mutex> g(mtx);
      nodes[ni].box = box;
      nodes[ni].child[0] = c0;
      nodes[ni].child[1] = c0 + 1;
    }
    uint32_t nl = mid - first, nr = first + count - mid;
    int dl, dr;
    if (count > 65536 && depthNow < 6) {
      auto fut = std::async(std::launch::async, [&]() { return build(c0, first, nl, depthNow + 1); });
      dr = build(c0 + 1, mid, nr, depthNow + 1);
      dl = fut.get();
    } else {
      dl = build(c0, first, nl, depthNow + 1);
      dr = build(c0 + 1, mid, nr, depthNow + 1);
    }
    return 1 + std::max(dl, dr);
  }
};

void padBox(const Box& b, float* lo, float* hi) {
  for (int c = 0; c < 3; ++c) {
    float m = std::max(std::fabs(b.lo[c]), std::fabs(b.hi[c]));
    float pad = m * 0x1p-18f + 1e-30f;
    lo[c] = b.lo[c] - pad;
    hi[c] = b.hi[c] + pad;
  }
}

}  // namespace

Bvh4 buildBvh4(const float* triPos, uint32_t numTris, const BvhBuildOptions& opt, int width) {
  if (width != 4 && width != 8) throw std::runtime_error("BVH width must be 4 or 8");
  Bvh4 out;
  out.width = width;
  out.numTris = numTris;
  const int W = width;
  const size_t nodeFloats = 8 * (size_t)W;
  if (numTris >= (1u << 24)) throw std::runtime_error("scene has more than 2^24 triangles");
  Builder B(opt);
  B.triBox.resize(numTris);
  B.cen.resize((size_t)numTris * 3);
  B.order.resize(numTris);
  for (uint32_t t = 0; t < numTris; ++t) {
    Box b;
    for (int v = 0; v < 3; ++v) b.growPoint(triPos + (size_t)t * 9 + v * 3);
    B.triBox[t] = b;
    for (int c = 0; c < 3; ++c) B.cen[(size_t)t * 3 + c] = 0.5f * (b.lo[c] + b.hi[c]);
    B.order[t] = t;
  }
  B.nodes.reserve(numTris > 0 ? 2 * (size_t)numTris : 1);
  B.nodes.emplace_back();
  if (numTris > 0) out.binaryDepth = B.build(0, 0, numTris, 0);
  out.binaryNodes = B.nodes.size();

  // Collapse to W-wide nodes, depth-first.
  const std::vector<BNode>& bn = B.nodes;
  std::vector<uint32_t> leafTris;
  leafTris.reserve(numTris);
  struct Work {
    int32_t bnode;
    uint32_t slot;  // index of the W-wide node to fill
    int depth;
  };
  std::vector<Work> stack;
  auto newNode = [&]() {
    uint32_t idx = out.numNodes++;
    out.nodes.resize((size_t)out.numNodes * nodeFloats, 0.0f);
    return idx;
  };
  // child k of node n: half k / 4 (32 floats), lane k % 4 in it
  auto writeChild = [&](uint32_t node, int k, const float* lo, const float* hi, uint32_t ref) {
    float* h = &out.nodes[(size_t)node * nodeFloats + 32 * (size_t)(k / 4)];
    const int l = k % 4;
    h[0 * 4 + l] = lo[0];
    h[1 * 4 + l] = hi[0];
    h[2 * 4 + l] = lo[1];
    h[3 * 4 + l] = hi[1];
    h[4 * 4 + l] = lo[2];
    h[5 * 4 + l] = hi[2];
    memcpy(&h[24 + l], &ref, 4);
  };
  const float emptyLo[3] = {INFINITY, INFINITY, INFINITY}, emptyHi[3] = {INFINITY, INFINITY, INFINITY};
  if (numTris == 0) {
    uint32_t r = newNode();
    for (int k = 0; k < W; ++k) writeChild(r, k, emptyLo, emptyHi, kLeafFlag);
    return out;
  }
  // The root of the 4-wide tree always is an inner node (a lone leaf becomes
  // its only child).
  stack.push_back({0, newNode(), 1});
  while (!stack.empty()) {
    Work w = stack.back();
    stack.pop_back();
    out.depth = std::max(out.depth, w.depth);
    std::vector<int32_t> kids;
    if (bn[w.bnode].count > 0) {
      kids.push_back(w.bnode);
    } else {
      kids.push_back(bn[w.bnode].child[0]);
      kids.push_back(bn[w.bnode].child[1]);
      while ((int)kids.size() < W) {
        int best = -1;
        float bestA = -1.0f;
        for (size_t i = 0; i < kids.size(); ++i) {
          const BNode& k = bn[kids[i]];
          if (k.count == 0 && k.box.area() > bestA) {
            bestA = k.box.area();
            best = (int)i;
          }
        }
        if (best < 0) break;
        int32_t b = kids[best];
        kids[best] = bn[b].child[0];
        kids.push_back(bn[b].child[1]);
      }
    }
    for (int k = 0; k < W; ++k) {
      if (k >= (int)kids.size()) {
        writeChild(w.slot, k, emptyLo, emptyHi, kLeafFlag);
        continue;
      }
      const BNode& kn = bn[kids[k]];
      float lo[3], hi[3];
      padBox(kn.box, lo, hi);
      uint32_t ref;
      if (kn.count > 0) {
        uint32_t firstOut = (uint32_t)leafTris.size();
        for (uint32_t i = 0; i < kn.count; ++i) leafTris.push_back(B.order[kn.first + i]);
        if (kn.count > 127) throw std::runtime_error("BVH leaf larger than 127 triangles");
        ref = makeLeafRef(firstOut, kn.count);
      } else {
        ref = newNode();
        stack.push_back({kids[k], ref, w.depth + 1});
      }
      writeChild(w.slot, k, lo, hi, ref);
    }
  }
  out.tris.resize((size_t)leafTris.size() * 12);
  for (size_t i = 0; i < leafTris.size(); ++i) {
    uint32_t t = leafTris[i];
    float* r = &out.tris[i * 12];
    for (int v = 0; v < 3; ++v)
      for (int c = 0; c < 3; ++c) r[v * 4 + c] = triPos[(size_t)t * 9 + v * 3 + c];
    memcpy(&r[3], &t, 4);
    r[7] = 0.0f;
    r[11] = 0.0f;
  }
  return out;
}

}  // namespace hasty

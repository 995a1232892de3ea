// ORACLE (test infrastructure only).  Closest-hit ray queries standing in for
// Embree's rtcIntersect1 (core/src/Scene.cpp:281-303): tnear = 1e-6,
// tfar = +inf, barycentrics (1-u-v, u, v), geometric normal along
// (v1-v0) x (v2-v0).
//
// Embree is not available here, so the oracle defines the hit itself:
//  * the watertight ray/triangle test of Woop, Benthin and Wald, "Watertight
//    Ray/Triangle Intersection", JCGT 2(1), 2013 (with its double-precision
//    fallback for exactly-zero edge functions);
//  * the closest hit is the smallest t, ties broken by the smaller global
//    triangle index, so the answer does not depend on traversal order;
//  * a canonical BVH2 (binned SAH, 16 bins, C_trav = C_isect = 1, at most 4
//    triangles per leaf, front-to-back traversal) whose node-visit and
//    triangle-test counts define the algorithmic bytes per ray (64 B per node,
//    48 B per triangle), plus brute force over every triangle.
// Boxes are padded and culling is relaxed so that the BVH never skips a
// triangle the test would report; both search orders must return the
// brute-force answer.
#pragma once
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cfloat>
#include <vector>

namespace oracle {

struct HitRecord {
  int32_t tri = -1;  // global triangle index, -1 = miss
  float t = 0.0f, u = 0.0f, v = 0.0f;
  bool edgeTie = false;  // an edge function of the hit was exactly zero
};

struct RayPre {
  float o[3], d[3];
  int kx, ky, kz;
  float Sx, Sy, Sz;
  float inv[3];
};

inline RayPre prepareRay(const float o[3], const float d[3]) {
  RayPre r;
  for (int i = 0; i < 3; ++i) {
    r.o[i] = o[i];
    r.d[i] = d[i];
    r.inv[i] = 1.0f / d[i];
  }
  float ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
  r.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  r.kx = r.kz + 1 == 3 ? 0 : r.kz + 1;
  r.ky = r.kx + 1 == 3 ? 0 : r.kx + 1;
  if (d[r.kz] < 0.0f) std::swap(r.kx, r.ky);
  r.Sx = d[r.kx] / d[r.kz];
  r.Sy = d[r.ky] / d[r.kz];
  r.Sz = 1.0f / d[r.kz];
  return r;
}

// Returns true and fills t/u/v if the ray hits the triangle with t > tnear.
inline bool intersectTriangle(const RayPre& r, const float* p, float tnear, float* tOut, float* uOut, float* vOut,
                              bool* edge) {
  float A[3], B[3], C[3];
  for (int i = 0; i < 3; ++i) {
    A[i] = p[i] - r.o[i];
    B[i] = p[3 + i] - r.o[i];
    C[i] = p[6 + i] - r.o[i];
  }
  float Ax = A[r.kx] - r.Sx * A[r.kz];
  float Ay = A[r.ky] - r.Sy * A[r.kz];
  float Bx = B[r.kx] - r.Sx * B[r.kz];
  float By = B[r.ky] - r.Sy * B[r.kz];
  float Cx = C[r.kx] - r.Sx * C[r.kz];
  float Cy = C[r.ky] - r.Sy * C[r.kz];
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  bool e = false;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    e = true;
    double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
    U = (float)(CxBy - CyBx);
    double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
    V = (float)(AxCy - AyCx);
    double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
    W = (float)(BxAy - ByAx);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float Az = r.Sz * A[r.kz];
  float Bz = r.Sz * B[r.kz];
  float Cz = r.Sz * C[r.kz];
  float T = U * Az + V * Bz + W * Cz;
  float t = T / det;
  if (!(t > tnear)) return false;
  *tOut = t;
  *uOut = V / det;
  *vOut = W / det;
  *edge = e && (U == 0.0f || V == 0.0f || W == 0.0f);
  return true;
}

inline bool closerHit(float t, int32_t tri, const HitRecord& best) {
  return best.tri < 0 || t < best.t || (t == best.t && tri < best.tri);
}

struct Bvh2 {
  struct Node {  // 64 bytes in the canonical layout
    float lo[3], hi[3];
    int32_t left = -1, right = -1;  // children (internal)
    int32_t first = 0, count = 0;   // triangle range (leaf when count > 0)
  };
  std::vector<Node> nodes;
  std::vector<uint32_t> triOrder;
  const float* triPos = nullptr;  // 9 floats per triangle
  uint32_t numTris = 0;

  struct Stats {
    uint64_t nodeVisits = 0, triTests = 0, rays = 0;
  };

  void build(const float* pos, uint32_t n) {
    triPos = pos;
    numTris = n;
    nodes.clear();
    triOrder.resize(n);
    for (uint32_t i = 0; i < n; ++i) triOrder[i] = i;
    if (n == 0) return;
    std::vector<float> cen((size_t)n * 3), blo((size_t)n * 3), bhi((size_t)n * 3);
    for (uint32_t i = 0; i < n; ++i) {
      for (int c = 0; c < 3; ++c) {
        float a = pos[i * 9 + c], b = pos[i * 9 + 3 + c], d = pos[i * 9 + 6 + c];
        float lo = std::min(a, std::min(b, d)), hi = std::max(a, std::max(b, d));
        blo[i * 3 + c] = lo;
        bhi[i * 3 + c] = hi;
        cen[i * 3 + c] = 0.5f * (lo + hi);
      }
    }
    nodes.reserve(2 * (size_t)n);
    nodes.push_back(Node());
    struct Item {
      int32_t node, first, count;
    };
    std::vector<Item> work{{0, 0, (int32_t)n}};
    while (!work.empty()) {
      Item it = work.back();
      work.pop_back();
      Node nd;
      float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
      for (int c = 0; c < 3; ++c) {
        nd.lo[c] = FLT_MAX;
        nd.hi[c] = -FLT_MAX;
      }
      for (int32_t k = it.first; k < it.first + it.count; ++k) {
        uint32_t t = triOrder[k];
        for (int c = 0; c < 3; ++c) {
          nd.lo[c] = std::min(nd.lo[c], blo[t * 3 + c]);
          nd.hi[c] = std::max(nd.hi[c], bhi[t * 3 + c]);
          clo[c] = std::min(clo[c], cen[t * 3 + c]);
          chi[c] = std::max(chi[c], cen[t * 3 + c]);
        }
      }
      // Pad by a few hundred ulps of the box magnitude.
      for (int c = 0; c < 3; ++c) {
        float m = std::max(std::fabs(nd.lo[c]), std::fabs(nd.hi[c]));
        float pad = m * 0x1p-18f + 1e-30f;
        nd.lo[c] -= pad;
        nd.hi[c] += pad;
      }
      if (it.count <= 4) {
        nd.first = it.first;
        nd.count = it.count;
        nodes[it.node] = nd;
        continue;
      }
      // Binned SAH over centroid bounds, 16 bins per axis.
      const int NB = 16;
      float bestCost = FLT_MAX;
      int bestAxis = -1, bestSplit = -1;
      for (int axis = 0; axis < 3; ++axis) {
        float ext = chi[axis] - clo[axis];
        if (!(ext > 0.0f)) continue;
        int cnt[NB] = {0};
        float bl[NB][3], bh[NB][3];
        for (int b = 0; b < NB; ++b)
          for (int c = 0; c < 3; ++c) {
            bl[b][c] = FLT_MAX;
            bh[b][c] = -FLT_MAX;
          }
        for (int32_t k = it.first; k < it.first + it.count; ++k) {
          uint32_t t = triOrder[k];
          int b = (int)((cen[t * 3 + axis] - clo[axis]) / ext * NB);
          b = std::min(std::max(b, 0), NB - 1);
          cnt[b]++;
          for (int c = 0; c < 3; ++c) {
            bl[b][c] = std::min(bl[b][c], blo[t * 3 + c]);
            bh[b][c] = std::max(bh[b][c], bhi[t * 3 + c]);
          }
        }
        auto area = [](const float* lo, const float* hi) {
          float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
          return 2.0f * (dx * dy + dy * dz + dz * dx);
        };
        float rl[NB][3], rh[NB][3];
        int rc[NB];
        float accl[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, acch[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        int accc = 0;
        for (int b = NB - 1; b > 0; --b) {
          accc += cnt[b];
          for (int c = 0; c < 3; ++c) {
            accl[c] = std::min(accl[c], bl[b][c]);
            acch[c] = std::max(acch[c], bh[b][c]);
          }
          rc[b] = accc;
          for (int c = 0; c < 3; ++c) {
            rl[b][c] = accl[c];
            rh[b][c] = acch[c];
          }
        }
        float ll[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, lh[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        int lc = 0;
        float parentArea = area(nd.lo, nd.hi);
        for (int b = 0; b < NB - 1; ++b) {
          lc += cnt[b];
          for (int c = 0; c < 3; ++c) {
            ll[c] = std::min(ll[c], bl[b][c]);
            lh[c] = std::max(lh[c], bh[b][c]);
          }
          int rcount = rc[b + 1];
          if (lc == 0 || rcount == 0) continue;
          float cost = 1.0f + (area(ll, lh) * lc + area(rl[b + 1], rh[b + 1]) * rcount) / parentArea;
          if (cost < bestCost) {
            bestCost = cost;
            bestAxis = axis;
            bestSplit = b;
          }
        }
      }
      int32_t mid;
      if (bestAxis < 0) {
        mid = it.first + it.count / 2;  // all centroids coincide
      } else {
        float ext = chi[bestAxis] - clo[bestAxis];
        auto part = std::partition(triOrder.begin() + it.first, triOrder.begin() + it.first + it.count,
                                   [&](uint32_t t) {
                                     int b = (int)((cen[t * 3 + bestAxis] - clo[bestAxis]) / ext * NB);
                                     b = std::min(std::max(b, 0), NB - 1);
                                     return b <= bestSplit;
                                   });
        mid = (int32_t)(part - triOrder.begin());
        if (mid == it.first || mid == it.first + it.count) mid = it.first + it.count / 2;
      }
      nd.left = (int32_t)nodes.size();
      nd.right = nd.left + 1;
      nodes.push_back(Node());
      nodes.push_back(Node());
      nodes[it.node] = nd;
      work.push_back({nd.right, mid, it.first + it.count - mid});
      work.push_back({nd.left, it.first, mid - it.first});
    }
  }

  // Conservative slab test (Ize, "Robust BVH Ray Traversal", JCGT 2013).
  static bool slab(const RayPre& r, const Node& n, float tbest, float* tEntry) {
    float tmin = -FLT_MAX, tmax = FLT_MAX;
    for (int c = 0; c < 3; ++c) {
      float t0 = (n.lo[c] - r.o[c]) * r.inv[c];
      float t1 = (n.hi[c] - r.o[c]) * r.inv[c];
      float a = std::fmin(t0, t1), b = std::fmax(t0, t1);
      if (a == a) tmin = std::fmax(tmin, a);
      if (b == b) tmax = std::fmin(tmax, b);
    }
    tmax *= 1.0f + 2.0f * 3.0f * 0x1p-24f / (1.0f - 3.0f * 0x1p-24f);
    *tEntry = tmin;
    return tmin <= tmax && tmax > 0.0f && tmin <= tbest * 1.001f;
  }

  // Front-to-back closest hit.  One node visit = one 64-byte record holding
  // both child boxes; the stack keeps each entry's entry distance so a popped
  // entry farther than the current best is dropped without a fetch.
  HitRecord closest(const float o[3], const float d[3], float tnear, Stats* stats) const {
    HitRecord best;
    best.t = FLT_MAX;
    if (nodes.empty()) return best;
    RayPre r = prepareRay(o, d);
    if (stats) stats->rays++;
    struct Entry {
      int32_t node;
      float t;
    };
    Entry stack[256];
    int sp = 0;
    float te;
    if (!slab(r, nodes[0], FLT_MAX, &te)) return best;
    int32_t cur = 0;
    while (true) {
      const Node& n = nodes[cur];
      if (n.count > 0) {
        for (int32_t k = n.first; k < n.first + n.count; ++k) {
          uint32_t tri = triOrder[k];
          if (stats) stats->triTests++;
          float t, u, v;
          bool e;
          if (intersectTriangle(r, triPos + (size_t)tri * 9, tnear, &t, &u, &v, &e) &&
              closerHit(t, (int32_t)tri, best)) {
            best.tri = (int32_t)tri;
            best.t = t;
            best.u = u;
            best.v = v;
            best.edgeTie = e;
          }
        }
      } else {
        if (stats) stats->nodeVisits++;
        float tb = best.tri < 0 ? FLT_MAX : best.t;
        float tl, tr;
        bool hl = slab(r, nodes[n.left], tb, &tl);
        bool hr = slab(r, nodes[n.right], tb, &tr);
        if (hl && hr) {
          bool leftFirst = tl <= tr;
          stack[sp++] = leftFirst ? Entry{n.right, tr} : Entry{n.left, tl};
          cur = leftFirst ? n.left : n.right;
          continue;
        }
        if (hl || hr) {
          cur = hl ? n.left : n.right;
          continue;
        }
      }
      bool found = false;
      while (sp > 0) {
        Entry en = stack[--sp];
        if (best.tri < 0 || en.t <= best.t * 1.001f) {
          cur = en.node;
          found = true;
          break;
        }
      }
      if (!found) break;
    }
    return best;
  }

  HitRecord bruteForce(const float o[3], const float d[3], float tnear) const {
    HitRecord best;
    RayPre r = prepareRay(o, d);
    for (uint32_t tri = 0; tri < numTris; ++tri) {
      float t, u, v;
      bool e;
      if (intersectTriangle(r, triPos + (size_t)tri * 9, tnear, &t, &u, &v, &e) && closerHit(t, (int32_t)tri, best)) {
        best.tri = (int32_t)tri;
        best.t = t;
        best.u = u;
        best.v = v;
        best.edgeTie = e;
      }
    }
    return best;
  }
};

}  // namespace oracle

// ORACLE (test infrastructure only).  CPU restatement of HastyRenderer's
// stochastic progressive photon mapper (caustic photon map + MIS path
// tracing): core/src/PhotonMapping.cpp, core/include/Hasty/PhotonMapping.h,
// core/include/Hasty/HashCells.h.
//
// Random streams (the reference consumes one shishua stream per thread):
//   initial radius of pixel p ........ (seed, p, 0, kDomainRadius)
//   photon i of pass k ............... (seed, i, k, kDomainPhoton)
//   camera path of pixel p, pass k ... (seed, p, k, kDomainPath)
// Photons of a pass are emitted in index order exactly like the reference's
// loop (PhotonMapping.cpp:861): photon i is emitted while fewer than
// `storeLimit` photons are stored and i < `emitLimit`.
#pragma once
#include <stdint.h>

#include <unordered_map>
#include <atomic>
#include <vector>

#include "oracle_pt.h"

namespace oracle {

struct Photon {  // PhotonMapping.cpp:11-23
  Vec3f wi;
  float ngwi;
  Vec3f x;
  Vec3f flux;
};

struct SppmParams {
  float alpha = 0.4f;           // PhotonMapping.h:65
  float footprintScale = 10.0f; // PhotonMapping.cpp:840
  double storeFraction = 0.5;   // countThreshold = 0.5 * W * H (PhotonMapping.cpp:834)
  double emitFraction = 10.0;   // tracethreshold = 10 * W * H (PhotonMapping.cpp:835)
  float initialRadius = 0.0f;   // > 0: this constant initial radius (no counterpart; PhotonMapping.cpp:838-841
                                // uses footprintScale x the pixel footprint)
  bool fixedPointFlux = false;  // gathers sum photon terms as int64 multiples of 2^-fluxFracBits
  int fluxFracBits = 24;
};

struct PixelStat {  // PMRenderJobPixel (PhotonMapping.h:12-31)
  float radius = -1.0f;
  float N = 0.0f;
  float Ne = 0.0f;
  Vec3f flux = Vec3f::Zero();
  Vec3f LSum = Vec3f::Zero();
  int Nd = 0;
};

struct CellKey {
  int32_t x, y, z;
  bool operator==(const CellKey& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct CellKeyHash {
  size_t operator()(const CellKey& k) const {
    uint64_t h = (uint64_t)(uint32_t)k.x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)k.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)(uint32_t)k.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

// PhotonMap with PointHashCells (PhotonMapping.cpp:71-180; HashCells.h:13-104)
struct PhotonMap {
  std::vector<Photon> photons;
  std::vector<Vec3f> points;
  size_t emitted = 0;
  float cellSize = 1.0f;
  // gathers in fixed point (no counterpart in the reference): terms rounded
  // to int64 units of 1 / fluxScale, saturated sums counted
  bool fixedFlux = false;
  float fluxScale = 0.0f;
  double fluxInvScale = 0.0;
  mutable std::atomic<uint32_t> fluxOverflows{0};
  std::unordered_map<CellKey, std::vector<uint32_t>, CellKeyHash> cells;

  void clear(float cs) {
    photons.clear();
    points.clear();
    emitted = 0;
    cellSize = cs;
    cells.clear();
  }
  CellKey index(const Vec3f& p) const {
    return CellKey{floor_int32(p[0] / cellSize), floor_int32(p[1] / cellSize), floor_int32(p[2] / cellSize)};
  }
  void build() {
    for (uint32_t i = 0; i < points.size(); ++i) cells[index(points[i])].push_back(i);
  }
  void radiusNeighbors(const Vec3f& c, float r, std::vector<uint32_t>& out) const;
};

struct SppmResult {  // PathTraceWithCausticsMapResult (PhotonMapping.cpp:394-403)
  Vec3f photonFlux = Vec3f::Zero();
  Vec3f LLight = Vec3f::Zero();
  Vec3f LRest = Vec3f::Zero();
  Vec3f LSpecular = Vec3f::Zero();
  Vec3f LCaustic = Vec3f::Zero();
  int M = 0;
};

struct SppmJob {
  PTJob pt;  // scene, settings, seed
  SppmParams params;
  std::vector<PixelStat> pixels;
  std::vector<float> radii;  // Image1f radii, row-major (x + y * W)
  bool initialised = false;
  uint32_t passesDone = 0;
  // last pass diagnostics
  uint64_t lastEmitted = 0, lastStored = 0;
  float lastCellSize = 0.0f;
  uint32_t lastFluxOverflows = 0;
};

void tracePhoton(RenderContext ctx, PhotonMap& map, const Ray& ray, const Vec3f& flux);
float computeCellSize(const std::vector<float>& radii, uint32_t W, uint32_t H);
float computeRadiusPixelFootPrint(RenderContext ctx, float defaultR, int i, int j, int W, int H);
float getCloseNonNegative(const std::vector<float>& img, int W, int H, int i, int j);
std::vector<float> computeInitialRadii(const PTJob& job, float footprintScale);

// One renderPhotonThread pass (PhotonMapping.cpp:848-929) for pass index
// `pass`: writes the current estimate into color (not sums), the depth-0
// normal and albedo into normal/albedo.
void sppmPass(SppmJob& job, uint32_t pass, float* color, float* normal, float* albedo);

}  // namespace oracle

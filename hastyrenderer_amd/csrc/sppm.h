// Stochastic progressive photon mapping on gfx950 (renderPhotonThread,
// core/src/PhotonMapping.cpp:814-933): host-side state and pass driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "pt_device.h"

namespace hasty {

struct SppmParams {
  float alpha = 0.4f;            // radius reduction (PhotonMapping.h:65)
  float footprintScale = 10.0f;  // initial radius = 10 x pixel footprint (PhotonMapping.cpp:840)
  double storeFraction = 0.5;    // stop emitting once 0.5 W H photons are stored (PhotonMapping.cpp:834)
  double emitFraction = 10.0;    // ... or 10 W H photons were emitted (PhotonMapping.cpp:835)
  uint32_t seed = 0;
};

struct SppmPassInfo {
  uint64_t emitted = 0, stored = 0, cells = 0;
  float cellSize = 0.0f;
  double traceMs = 0, buildMs = 0, gatherMs = 0;
};

class SppmRenderer {
 public:
  SppmRenderer(const hdev::DevScene& scene, const SppmParams& params, int bvhDepth, hipStream_t stream);
  ~SppmRenderer();
  // One pass: photon map, camera paths with gathers, statistics update.
  // Writes the current estimate (not a sum) and the depth-0 normal/albedo
  // into the device films.
  void pass(uint32_t passIndex, float* dColor, float* dNormal, float* dAlbedo);
  void pixelStats(float* host10);  // radius, N, Ne, flux rgb, LSum rgb, Nd
  void radii(float* host);
  const SppmPassInfo& lastInfo() const { return info_; }
  uint32_t passesDone() const { return passes_; }

 private:
  void init();
  void ensurePhotonCapacity(size_t n);
  hdev::DevScene s_;
  SppmParams p_;
  hipStream_t stream_;
  uint32_t W_, H_;
  bool initialised_ = false;
  uint32_t passes_ = 0;
  SppmPassInfo info_;
  std::vector<void*> allocs_;
  template <class T>
  T* alloc(size_t n);
  // per pixel
  float* dRadii_ = nullptr;
  float* dPixel_ = nullptr;  // 10 floats per pixel
  // photon batches
  uint32_t batch_ = 0;
  float4* dSlots_ = nullptr;  // per-thread slots: batch x maxDepth x 3 float4
  uint32_t* dSlotCount_ = nullptr;
  uint32_t* dSlotOffset_ = nullptr;
  void* dScanTmp_ = nullptr;
  size_t scanTmpBytes_ = 0;
  // photons of the pass (emission order), sorted copy by cell
  size_t photonCap_ = 0;
  float4* dPhotons_ = nullptr;        // 3 float4 per photon
  float4* dSorted_ = nullptr;         // 3 float4 per photon, cell order
  unsigned long long* dKeys_ = nullptr;
  unsigned long long* dKeysSorted_ = nullptr;
  uint32_t* dIdx_ = nullptr;
  uint32_t* dIdxSorted_ = nullptr;
  void* dSortTmp_ = nullptr;
  size_t sortTmpBytes_ = 0;
  // hash table of cells
  uint32_t tableSize_ = 0;
  unsigned long long* dTableKeys_ = nullptr;
  uint2* dTableVals_ = nullptr;
  uint32_t* dCounters_ = nullptr;
  float* dCellSize_ = nullptr;
  uint2* dOverflow_ = nullptr;
  uint32_t traceThreads_ = 0;
  std::vector<void*> tmpAllocs_;
};

}  // namespace hasty

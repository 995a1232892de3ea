// Stochastic progressive photon mapping on gfx950 (renderPhotonThread,
// core/src/PhotonMapping.cpp:814-933): host-side state and pass driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "pt_device.h"
#include "pt_kernels.h"

namespace hasty {

struct SppmParams {
  float alpha = 0.4f;            // radius reduction (PhotonMapping.h:65)
  float footprintScale = 10.0f;  // initial radius = 10 x pixel footprint (PhotonMapping.cpp:840)
  double storeFraction = 0.5;    // stop emitting once 0.5 W H photons are stored (PhotonMapping.cpp:834)
  double emitFraction = 10.0;    // ... or 10 W H photons were emitted (PhotonMapping.cpp:835)
  float initialRadius = 0.0f;    // > 0: this constant initial radius instead of the footprint estimate
  bool fixedPointFlux = false;   // gathers sum photon terms in int64 units of 2^-fluxFracBits
  int fluxFracBits = 24;
  uint32_t gatherShareCandidates = 0;  // 0: default threshold for a wave's shared gather
  float gatherUnionFactor = 0.0f;      // > 0: union gathers (see kSppmGather); 0: off
};

struct SppmPassInfo {
  uint64_t emitted = 0, stored = 0;
  float cellSize = 0.0f;
  uint32_t fluxOverflows = 0;  // gathers whose fixed-point sum saturated
  uint64_t gatherQueries = 0, gatherCandidates = 0, gatherFound = 0;
};

// Device state of one progressive photon-mapping job: per-pixel statistics
// (PMRenderJobPixel), the radius image and the per-pass photon map.
class SppmRenderer {
 public:
  SppmRenderer(const hdev::DevScene& scene, const SppmParams& params, hipStream_t stream);
  ~SppmRenderer();
  SppmRenderer(const SppmRenderer&) = delete;
  SppmRenderer& operator=(const SppmRenderer&) = delete;
  // One pass with pixel offset (offX, offY): photon map, camera paths with
  // gathers, statistics update.  Writes the current estimate (not a sum) and
  // the depth-0 normal / albedo into the device films.
  void pass(uint32_t passIndex, float offX, float offY, float* dColor, float* dNormal, float* dAlbedo,
            KernelTimer* timer);
  // radius, N, Ne, flux rgb, LSum rgb, Nd (int bits): 10 planes of W*H floats
  void pixelStats(float* host10);
  void radii(float* host);
  // photons of the last pass in emission order: x, ngwi, wi, flux (10 floats)
  void photons(float* host, size_t maxPhotons);
  const SppmPassInfo& lastInfo() const { return info_; }
  uint32_t passesDone() const { return passes_; }
  uint32_t emitLimit() const { return emitLimit_; }
  uint32_t storeLimit() const { return storeLimit_; }

 private:
  void initRadii();
  template <class T>
  T* alloc(size_t n);

  hdev::DevScene s_;
  SppmParams p_;
  hipStream_t stream_;
  uint32_t W_ = 0, H_ = 0;
  bool initialised_ = false;
  uint32_t passes_ = 0;
  SppmPassInfo info_;
  std::vector<void*> allocs_;
  uint32_t traceBlocks_ = 0, partialBlocks_ = 0;
  uint32_t emitLimit_ = 0, storeLimit_ = 0;
  // per pixel
  float* dPixel_ = nullptr;  // 10 planes
  float* dRadii_ = nullptr;
  // gather queries of the pass's camera paths (maxDepth x P records of 64
  // bytes), queries per pixel, path radiance (3 planes)
  float4* dQuery_ = nullptr;
  uint32_t* dNq_ = nullptr;
  float* dLpath_ = nullptr;
  // cell size reductions
  float* dPartial_ = nullptr;
  float* dCand_ = nullptr;
  float* dState_ = nullptr;
  float* dCellSize_ = nullptr;
  int* dBox_ = nullptr;
  uint32_t* dCounters_ = nullptr;
  unsigned long long* dGatherStats_ = nullptr;
  uint2* dOverflow_ = nullptr;
  // photon batches: per-thread slots, batch x cap
  uint32_t cap_ = 0, batchMax_ = 0;
  float4* dSlX_ = nullptr;
  float4* dSlWi_ = nullptr;
  float4* dSlF_ = nullptr;
  uint32_t* dSlCount_ = nullptr;
  uint32_t* dSlOffset_ = nullptr;
  void* dScanTmp_ = nullptr;
  size_t scanTmpBytes_ = 0;
  // photons of the pass in emission order and in cell order
  size_t photonCap_ = 0;
  float4* dPhX_ = nullptr;
  float4* dPhWi_ = nullptr;
  float4* dPhF_ = nullptr;
  float4* dSX_ = nullptr;
  float4* dSWi_ = nullptr;
  float4* dSF_ = nullptr;
  unsigned long long* dKeys_ = nullptr;
  unsigned long long* dKeysSorted_ = nullptr;
  uint32_t* dIdx_ = nullptr;
  uint32_t* dIdxSorted_ = nullptr;
  void* dSortTmp_ = nullptr;
  size_t sortTmpBytes_ = 0;
  // cell table
  uint32_t tableSize_ = 0;
  unsigned long long* dTableKeys_ = nullptr;
  uint2* dTableVals_ = nullptr;
};

}  // namespace hasty

// Wavefront path-tracing kernels: host-side launch interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_device.h"

namespace hasty {

// Structure-of-arrays path state for `capacity` paths (K samples x padded
// pixels).  Each array is indexed by path slot; queues hold slot indices.
struct PathBuffers {
  float4* rayO;   // origin; w unused
  float4* rayD;   // direction; w = pdf of the BSDF sample that produced it
  float4* hit;    // t, u, v, triangle (bits)
  float4* beta;   // path throughput
  float4* L;      // accumulated radiance
  uint4* state;   // x rng draws, y flags, z previous triangle, w unused
  uint4* media;   // 2 x uint4 per slot (8 entries)
  float4* auxN;   // depth-0 shading normal
  float4* auxA;   // depth-0 albedo
  float4* shO;    // shadow origin; w = BSDF pdf of the light direction
  float4* shD;    // shadow direction; w = target triangle (bits, -1 = environment)
  float4* shFD;   // f_diffuse; w = |cos| at the shading point
  float4* shFC;   // f_concentrated; w = shading triangle (bits)
  float4* shThr;  // local throughput (Beer / (1 - q))
  float4* shBeta; // path throughput at the vertex
  float4* shT;    // transmittance of the light-ray media copy
  uint32_t* qA;
  uint32_t* qB;
  uint32_t* qS;
  uint32_t* counters;  // [0] qA size [1] qB size [2] shadow size
  unsigned long long* stats;  // see StatId
  uint2* stackOverflow;
  uint32_t capacity;
  uint32_t traceThreads;  // threads of the traversal grids (overflow slices)
};

constexpr int kStackLds = 16;
constexpr int kStackOvf = 64;
constexpr int kBlock = 256;

constexpr int kNumCounters = 4;
// 64-bit device statistics (filled when LaunchConfig::countStats is set).
enum StatId { kStatExtRays = 0, kStatShadowRays, kStatNodes, kStatTris, kStatMediaOverflow, kNumStats = 8 };

struct LaunchConfig {
  uint32_t traceBlocks = 2048;  // persistent-style grid of the traversal kernels
  uint32_t shadeBlocks = 4096;
  bool countStats = false;
};

enum KernelId {
  kKGenerate = 0,
  kKExtend,
  kKShade,
  kKConnect,
  kKAccumulate,
  kKMisc,
  kKSppmRadii,
  kKSppmCellSize,
  kKSppmPhotons,
  kKSppmBuild,
  kKSppmCamera,
  kNumKernels
};
const char* kernelName(int id);

// Timing hook: called around each launch when profiling is on.
struct KernelTimer {
  virtual void begin(int kernel, hipStream_t s) = 0;
  virtual void end(int kernel, hipStream_t s) = 0;
  virtual ~KernelTimer() {}
};

// Renders samples [firstSample, firstSample + K) of every pixel and adds them,
// in sample order, into the device films (row-major RGB float32).
void launchPTBatch(const hdev::DevScene& scene, const PathBuffers& pb, uint32_t firstSample, uint32_t K,
                   const float2* dOffsets, float* dColor, float* dNormal, float* dAlbedo, const LaunchConfig& cfg,
                   hipStream_t stream, KernelTimer* timer);

// Closest hit for n rays (6 floats each): tri, t, u, v.
void launchTraceClosest(const hdev::DevScene& scene, const float* dRays, uint32_t n, int32_t* dTri, float* dTuv,
                        uint2* dOverflow, uint32_t overflowThreads, hipStream_t stream);

// Slots per sample: pixels padded to whole 8x8 tiles.
__host__ __device__ inline uint32_t paddedPixels(uint32_t width, uint32_t height) {
  return ((width + 7) / 8) * ((height + 7) / 8) * 64;
}

}  // namespace hasty

// Wavefront path-tracing kernels: host-side launch interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "guide.h"
#include "pt_device.h"

#include <vector>

namespace hasty {

// Structure-of-arrays path state for `capacity` path slots.  Slots are
// recycled: when a path ends its radiance goes to the results ring and the
// slot starts the next path of the render (regeneration), so every stage
// keeps a full queue until the last paths of a render drain.  Queues hold
// slot indices.
struct PathBuffers {
  float4* rayO;   // origin; w unused
  float4* rayD;   // direction; w = pdf of the BSDF sample that produced it
  float4* hit;    // t, u, v, triangle (bits)
  float4* beta;   // path throughput
  float4* L;      // accumulated radiance
  uint4* state;   // x rng draws, y flags, z previous triangle, w path id
  uint4* media;   // 2 x uint4 per slot (8 entries)
  uint32_t* mediaTx;  // 8 per slot: albedo texel of each media entry
  float4* auxN;   // depth-0 shading normal
  float4* auxA;   // depth-0 albedo
  float4* shO;    // shadow origin; w = BSDF pdf of the light direction
  float4* shD;    // shadow direction; w = target triangle (bits, -1 = environment)
  float4* shFD;   // f_diffuse; w = |cos| at the shading point
  float4* shFC;   // f_concentrated; w = shading triangle (bits)
  float4* shThr;  // local throughput (Beer / (1 - q))
  float4* shBeta; // path throughput at the vertex
  float4* shT;    // transmittance of the light-ray media copy
  uint32_t* qA;   // extension queue of this bounce
  uint32_t* qB;   // extension queue of the next bounce
  uint32_t* qS;   // shadow queue
  uint32_t* qR;   // paths that ended at this bounce
  uint32_t* mark; // bounce (epoch) whose extension queue holds the slot
  uint32_t* counters;         // see CounterId
  unsigned long long* stats;  // see StatId
  uint2* stackOverflow;
  float4* ringL;  // finished paths: radiance, normal, albedo at
  float4* ringN;  // ring slot (global path id % ringPaths)
  float4* ringA;
  uint32_t ringPaths;
  uint32_t capacity;
  uint32_t traceThreads;  // threads of the traversal grids (overflow slices)
};

enum CounterId {
  kCntA = 0,      // live paths in qA
  kCntB = 1,      // paths queued for the next bounce
  kCntS = 2,      // shadow rays
  kCntR = 3,      // paths that ended at this bounce
  kCntNext = 4,   // next path id to start
  kCntLimit = 5,  // path ids below this may start (results ring space)
  kCntFill = 6,   // idle slots seen by kFill this bounce (its cursor)
  kCntEpoch = 7,  // bounces advanced since the render started
  kNumCounters = 8
};

// One job of a render call: samples first .. first + paths / pixels - 1 of
// every pixel of its scene.  Path ids are global over the jobs of a call:
// job j owns [pidBase, pidBase + paths); inside a job, local id =
// (sample - first) * pixels + rank with pixel = tileOrder[rank] (8x8 tiles).
// Finished paths go to ring slot (global id % ringPaths).
struct PassInfo {
  uint32_t first;
  uint32_t pixels;
  uint32_t paths;
  uint32_t pidBase;
  const uint32_t* tileOrder;  // rank -> pixel
  const uint32_t* rankOf;     // pixel -> rank
  const float2* offsets;      // per sample of the call
};

// Guided rendering (single job): per-slot raw guide outputs, per-slot
// training records (2 x {position, normal, wo + the density wi was drawn
// with, wi + the BSDF density of wi, throughput before the sample, radiance
// before the continuation}) and the training samples harvested when paths
// end (4 float4 each: position + weight, normal, wo + drawing density,
// wi + BSDF density).
struct GuideRun {
  guide::Net net;
  float4* raw;      // 10 float4 per slot
  float4* rec;      // 12 float4 per slot
  float4* samples;  // sampleCap x 4 float4
  uint32_t* sampleCount;
  uint32_t sampleCap;
  float fraction;   // probability of drawing the direction from the guide
  float* sel;       // per slot: that probability as the guide learned it (or null: `fraction`)
  uint32_t maxDepth;  // only vertices of depth below this are guided
  float maxThroughput;  // ... and only while every channel of the path throughput is at most this
};

// The jobs of a render call.  With one job its scene and pass travel in the
// kernel arguments (`s`, `pi`); with several, kernels look them up by path id.
struct JobSet {
  hdev::DevScene s;
  PassInfo pi;
  const hdev::DevScene* scenes;
  const PassInfo* passes;
  uint32_t nJobs;
  uint32_t paths;  // over all jobs
  GuideRun g;      // read by the guided kernels only
};

// Traversal stack entries kept in LDS per lane (the rest spill to a global
// slice); a build knob because it trades LDS (occupancy) for spills.
#ifndef HASTY_STACK_LDS
#define HASTY_STACK_LDS 8
#endif
constexpr int kStackLds = HASTY_STACK_LDS;
constexpr int kStackOvf = 80 - kStackLds;  // 80 entries per lane in all
constexpr int kBlock = 256;

// 64-bit device statistics (filled when LaunchConfig::countStats is set).
// kStatNodes / kStatTris count the extension rays' node visits and triangle
// tests, kStatShadowNodes / kStatShadowTris the shadow rays'.
enum StatId {
  kStatExtRays = 0,
  kStatShadowRays,
  kStatNodes,
  kStatTris,
  kStatMediaOverflow,
  kStatShadowNodes,
  kStatShadowTris,
  kNumStats = 8
};

struct LaunchConfig {
  uint32_t traceBlocks = 2048;  // persistent-style grid of the traversal kernels
  uint32_t shadeBlocks = 8192;
  uint32_t chunk = 8;           // bounces between host checks of the live count
  bool useGraph = false;        // replay each chunk of bounces as one captured hipGraph
  bool countStats = false;
};

enum KernelId {
  kKGenerate = 0,
  kKExtend,
  kKShade,
  kKConnect,
  kKAccumulate,
  kKMisc,
  kKRetire,
  kKGraph,
  kKGuideInfer,
  kKSppmRadii,
  kKSppmCellSize,
  kKSppmPhotons,
  kKSppmBuild,
  kKSppmCamera,
  kKSppmGather,
  kNumKernels
};
const char* kernelName(int id);

// Timing hook: called around each launch when profiling is on.
struct KernelTimer {
  virtual void begin(int kernel, hipStream_t s) = 0;
  virtual void end(int kernel, hipStream_t s) = 0;
  virtual ~KernelTimer() {}
};

// One job of renderJobs: scene (seed set), samples [first, first + n), the n
// pixel offsets, the tile order and its inverse, device films (sums, added to
// in sample order).
struct RenderJobDesc {
  hdev::DevScene scene;
  uint32_t first, n;
  const float2* dOffsets;
  const uint32_t* dTileOrder;
  const uint32_t* dRankOf;
  float* dColor;
  float* dNormal;
  float* dAlbedo;
};

// Device bytes of the scene and pass tables of an n-job render call.
size_t jobTableBytes(uint32_t nJobs);

// Renders every job through one shared pool of path slots (paths of all jobs
// in flight together; at most 2^31 paths per call).  Synchronises with the
// stream once per cfg.chunk bounces.  With more than one job the per-job
// tables go to `jobTables` (device memory of jobTableBytes(n) bytes at least).
void renderJobs(const std::vector<RenderJobDesc>& jobs, const PathBuffers& pb, const LaunchConfig& cfg,
                hipStream_t stream, KernelTimer* timer, const GuideRun* guide = nullptr, void* jobTables = nullptr);

// Pixels of a width x height film in 8x8-tile order (row-major inside a
// tile), and the inverse map.
void tileOrder(uint32_t width, uint32_t height, uint32_t* order, uint32_t* rankOf);

// Hit record of a batched closest-hit query: the layout of HastyHit
// (include/hasty_amd.h): geomID, primID (0xFFFFFFFF on a miss), t, u, v,
// unit geometric normal.
struct HitRecord {
  uint32_t geomID, primID;
  float t, u, v;
  float ng[3];
};
static_assert(sizeof(HitRecord) == 32, "HitRecord is 32 bytes");

// Closest-hit records for n rays (6 floats each).  `overflow` holds
// kStackOvf stack entries for each of the overflowThreads threads the grid
// may use (a multiple of kBlock).
void launchTraceHits(const hdev::DevScene& scene, const float* dRays, uint32_t n, HitRecord* dHits, uint2* dOverflow,
                     uint32_t overflowThreads, hipStream_t stream);

// Caustic photons as guide training samples (4 float4 each, appended at
// atomicAdd(sampleCount) below cap) from `count` light paths of pass `pass`;
// `overflow` as for launchTraceHits.  Nothing without surface lights.
void launchGuidePhotons(const hdev::DevScene& scene, uint32_t pass, uint32_t count, uint2* dOverflow,
                        uint32_t overflowThreads, float4* samples, uint32_t* sampleCount, uint32_t cap,
                        hipStream_t stream);

// Closest hit for n rays (6 floats each): tri, t, u, v.
void launchTraceClosest(const hdev::DevScene& scene, const float* dRays, uint32_t n, int32_t* dTri, float* dTuv,
                        uint2* dOverflow, uint32_t overflowThreads, hipStream_t stream);


}  // namespace hasty

// hasty_trace_hits_device_noalloc: the closest-hit query of
// hasty_trace_hits_device for callers that cannot allocate device memory.
//
// The only device memory the query needs beyond its rays and hits is the
// traversal stack's overflow part (stack entries past the kStackLds kept in
// LDS).  hasty_trace_hits_device takes it from a buffer the scene allocates on
// first use; here it is `scratch`, a file-scope device array sized for a
// fixed grid of kNoAllocBlocks blocks (9 MiB), which every GPU the library runs
// on holds until the process ends.  Its device address is looked up once per
// device with hipGetSymbolAddress (the array's name is only a host-side
// stand-in) and passed to the same kernel as hasty_trace_hits_device uses.
// Stack slices are written before they are read, so nothing needs clearing
// between calls.  Calls that overlap on one GPU (two host threads, or scenes
// on different streams) share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/hasty_amd.h"
#include "capi_internal.h"
#include "pt_kernels.h"

namespace {
constexpr uint32_t kNoAllocBlocks = 64;
constexpr uint32_t kNoAllocThreads = kNoAllocBlocks * hasty::kBlock;
constexpr size_t kScratchEntries = (size_t)kNoAllocThreads * hasty::kStackOvf;  // 9 MiB of uint2
}  // namespace

static __device__ __align__(256) uint2 scratch[kScratchEntries];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
uint2* g_scratchPtr[kMaxDevices] = {};

void hipOk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device address of `scratch` on `device` (the current device), looked up
// on the first call on that device.
uint2* scratchOn(int device) {
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range");
  if (!g_scratchPtr[device]) {
    void* p = nullptr;
    hipOk(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
    g_scratchPtr[device] = static_cast<uint2*>(p);
  }
  return g_scratchPtr[device];
}

}  // namespace

extern "C" int hasty_trace_hits_device_noalloc(HastyScene* scene, const HastyRay* d_rays, uint32_t n,
                                               HastyHit* d_hits) {
  hasty::SceneView v;
  if (!hasty::sceneView(scene, &v) || (n > 0 && (!d_rays || !d_hits))) {
    hasty::setLastError("null argument");
    return 1;
  }
  if (((uintptr_t)d_hits & 15u) != 0 || ((uintptr_t)d_rays & 3u) != 0) {
    hasty::setLastError("hits must be 16-byte aligned and rays 4-byte aligned");
    return 1;
  }
  if (n >= 0x80000000u) {
    hasty::setLastError("more than 2^31 - 1 rays in one query (the no-allocation query's maximum)");
    return 1;
  }
  // A lane's stack slice holds kStackLds + kStackOvf entries; the traversal
  // pushes at most width - 1 per inner level plus the root.
  if ((v.bvhWidth - 1) * v.bvhDepth + 1 > hasty::kStackLds + hasty::kStackOvf) {
    hasty::setLastError("BVH too deep for the no-allocation traversal stack (depth " + std::to_string(v.bvhDepth) +
                        ")");
    return 1;
  }
  if (n == 0) return 0;
  try {
    hipOk(hipSetDevice(v.device), "hipSetDevice");
    uint2* ovf = scratchOn(v.device);
    hasty::launchTraceHits(v.dev, reinterpret_cast<const float*>(d_rays), n,
                           reinterpret_cast<hasty::HitRecord*>(d_hits), ovf, kNoAllocThreads, v.stream);
    hipOk(hipGetLastError(), "kTraceHits launch");
    hipOk(hipStreamSynchronize(v.stream), "hipStreamSynchronize");
    return 0;
  } catch (const std::exception& e) {
    hasty::setLastError(e.what());
    return 3;
  }
}

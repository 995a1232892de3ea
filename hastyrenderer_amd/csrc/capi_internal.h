// Internal view of a HastyScene for entry points compiled in other
// translation units of the library (not part of the public ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/hasty_amd.h"
#include "pt_device.h"

namespace hasty {

struct SceneView {
  hdev::DevScene dev;  // device pointers and constants of the loaded scene
  hipStream_t stream;  // the scene's stream
  int device;          // the device the scene lives on
  int bvhDepth;        // inner levels of the BVH in use
  int bvhWidth;        // its children per node (4 or 8)
};

// Fills *out; false for a null scene.
bool sceneView(const HastyScene* scene, SceneView* out);

// Sets the message hasty_last_error() returns on the calling thread.
void setLastError(const std::string& message);

// computeEstimate of `count` device films of `floats` values each into `out`
// on `stream` (asynchronous): counts[k] = samples in film k, scale = the
// exposure factor (output_kernels.hip).
void launchEstimate(const float* const* films, const uint32_t* counts, int count, float scale, size_t floats,
                    float* out, hipStream_t stream);

}  // namespace hasty

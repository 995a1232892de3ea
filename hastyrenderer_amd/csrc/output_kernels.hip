// The output step after the path, on device films:
//   hasty_estimate_device  computeEstimate (apps/ViewerUtils.cpp:95-121) with
//                          the viewers' exposure scale (PathTracingViewer.cpp:
//                          120, 160)
//   hasty_present_device   preparePresentableImage (ViewerUtils.cpp:15-65):
//                          pow(clamp(v, 0, 1), 1 / 2.2) and RGBA8 packing
// Both are elementwise passes over W*H*3 floats (HBM-bound: 12 B read per
// film and 12 B written per pixel for the estimate; 12 B read and 4 B written
// per pixel for the presentable image).
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>

#include "../../include/hasty_amd.h"
#include "capi_internal.h"
#include "hmath.h"

namespace {

constexpr int kMaxFilms = 64;
constexpr int kThreads = 256;

struct FilmList {
  const float* film[kMaxFilms];
  int count;
  float denominator;  // 1e-7 + total samples
  float scale;        // 2^exposure (1 for normal / albedo films)
};

// image = films[0] + films[1] + ... (left to right), / denominator, * scale
__global__ void __launch_bounds__(kThreads) kEstimate(FilmList f, size_t n, float* __restrict__ out) {
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
    float v = f.film[0][i];
    for (int k = 1; k < f.count; ++k) v += f.film[k][i];
    v /= f.denominator;
    if (f.scale != 1.0f) v *= f.scale;
    out[i] = v;
  }
}

__device__ inline uint32_t channel8(float c) {
  float g = c > 0.0f ? (c < 1.0f ? hm::powf_pos(c, 1.0f / 2.2f) : 1.0f) : 0.0f;  // clampAndPow(0, 1, 1 / 2.2)
  return (uint32_t)(255.0f * g) & 0xFFu;                                          // toUint32: 255 * clamp(c, 0, 1)
}
__global__ void __launch_bounds__(kThreads) kPresent(const float* __restrict__ img, size_t pixels,
                                                     uint32_t* __restrict__ rgba) {
  for (size_t p = (size_t)blockIdx.x * kThreads + threadIdx.x; p < pixels; p += (size_t)gridDim.x * kThreads) {
    const float* c = img + 3 * p;
    rgba[p] = channel8(c[0]) << 24 | channel8(c[1]) << 16 | channel8(c[2]) << 8 | 0xFFu;
  }
}

void hipOk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

unsigned gridFor(size_t n) {
  size_t blocks = (n + kThreads - 1) / kThreads;
  return (unsigned)std::min<size_t>(std::max<size_t>(blocks, 1), 8192);
}

}  // namespace

namespace hasty {
// The estimate of `count` films of `floats` values each into `out` on
// `stream`; counts[k] = samples in film k (computeEstimate's totals).
void launchEstimate(const float* const* films, const uint32_t* counts, int count, float scale, size_t floats,
                    float* out, hipStream_t stream) {
  if (count < 1 || count > kMaxFilms) throw std::runtime_error("1 to 64 films per estimate");
  // computeEstimate replaces the image with each buffer until the running
  // sample total is nonzero, then adds: the sum starts at the first film
  // with samples (or is the last film when none has any)
  int start = count - 1;
  for (int k = 0; k < count; ++k)
    if (counts[k] > 0) {
      start = k;
      break;
    }
  FilmList f{};
  int total = 0;
  for (int k = 0; k < count; ++k) total += (int)counts[k];
  for (int k = start; k < count; ++k) {
    if (!films[k]) throw std::runtime_error("null film");
    f.film[k - start] = films[k];
  }
  f.count = count - start;
  f.denominator = 1e-7f + float(total);
  f.scale = scale;
  hipLaunchKernelGGL(kEstimate, dim3(gridFor(floats)), dim3(kThreads), 0, stream, f, floats, out);
  hipOk(hipGetLastError(), "kEstimate launch");
}
}  // namespace hasty

extern "C" int hasty_estimate_device(HastyScene* scene, const float* const* d_films, const uint32_t* num_samples,
                                     uint32_t n_films, float exposure, float* d_out) {
  hasty::SceneView v;
  if (!hasty::sceneView(scene, &v) || !d_films || !num_samples || !d_out) {
    hasty::setLastError("null argument");
    return 1;
  }
  if (n_films < 1 || n_films > (uint32_t)kMaxFilms) {
    hasty::setLastError("1 to 64 films per estimate");
    return 1;
  }
  try {
    hipOk(hipSetDevice(v.device), "hipSetDevice");
    const float scale = exposure == 0.0f ? 1.0f : std::pow(2.0f, exposure);  // PathTracingViewer.cpp:120
    hasty::launchEstimate(d_films, num_samples, (int)n_films, scale, (size_t)v.dev.width * v.dev.height * 3, d_out,
                          v.stream);
    hipOk(hipStreamSynchronize(v.stream), "hipStreamSynchronize");
    return 0;
  } catch (const std::exception& e) {
    hasty::setLastError(e.what());
    return 3;
  }
}

extern "C" int hasty_present_device(HastyScene* scene, const float* d_image, uint32_t* d_rgba) {
  hasty::SceneView v;
  if (!hasty::sceneView(scene, &v) || !d_image || !d_rgba) {
    hasty::setLastError("null argument");
    return 1;
  }
  try {
    hipOk(hipSetDevice(v.device), "hipSetDevice");
    const size_t pixels = (size_t)v.dev.width * v.dev.height;
    hipLaunchKernelGGL(kPresent, dim3(gridFor(pixels)), dim3(kThreads), 0, v.stream, d_image, pixels, d_rgba);
    hipOk(hipGetLastError(), "kPresent launch");
    hipOk(hipStreamSynchronize(v.stream), "hipStreamSynchronize");
    return 0;
  } catch (const std::exception& e) {
    hasty::setLastError(e.what());
    return 3;
  }
}

// C++ host layer over the C ABI (include/hasty_amd.h) with the reference's
// render-thread interface, so that the reference's apps can call the GPU
// renderer through the same names and signatures:
//   renderThread        core/include/Hasty/PathTracing.h:63, PathTracing.cpp:98-160
//   renderPhotonThread  core/include/Hasty/PhotonMapping.h:124, PhotonMapping.cpp:814-933
//   RenderJob           core/include/Hasty/RenderJob.h:42-64
//   RenderThreadData    core/include/Hasty/PathTracing.h:15-26
//   DoubleBuffer / AccumulationBuffer  core/include/Hasty/Image.h:18-45, 302-316
//   computeEstimate     apps/ViewerUtils.cpp:95-121
//   writeEXR            core/src/Image.cpp:140-152
// Images are row-major RGB float32 (W*H*3), row 0 at the top.
#pragma once
#include <stdint.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hasty_amd.h"

namespace Hasty {

// RenderSettings (RenderJob.h:15-30) as read from the job file.
struct RenderSettings {
  int rouletteStartDepth = 1;
  float rouletteQ = 0.3f;
  uint32_t width = 720;
  uint32_t height = 720;
  uint32_t numThreads = 4;
  int maxDepth = 30;
  int numSamples = 32;
  bool useMIS = true;
  float exposure = 0.0f;
};

struct RenderJobSample {
  int idx;
};

// The GPU scene of a job; calls on it are serialised by its mutex.
class GpuScene {
 public:
  explicit GpuScene(const std::string& renderJobPath);  // throws std::runtime_error
  ~GpuScene();
  GpuScene(const GpuScene&) = delete;
  GpuScene& operator=(const GpuScene&) = delete;
  HastyScene* handle() { return h_; }
  std::mutex& mutex() { return m_; }

 private:
  HastyScene* h_ = nullptr;
  std::mutex m_;
};

class RenderJob {
 public:
  virtual ~RenderJob() {}
  void loadJSON(const std::string& jsonFilepath);  // throws std::runtime_error like the reference
  RenderSettings renderSettings;
  float cameraExposure = 0.0f;  // Camera::getExposure (Camera.h:34), the viewers' display scale 2^exposure
  std::shared_ptr<GpuScene> scene;
  std::atomic<int> sampleIdxGenerator{0};
  std::atomic<bool> stopFlag{false};
  RenderJobSample fetchSample() {
    std::lock_guard<std::mutex> lock(m_sampleMutex);
    return RenderJobSample{sampleIdxGenerator++};
  }
  // A contiguous run of up to `count` sample indices below numSamples
  // (first index, number); number 0 when the job is exhausted.
  void fetchSamples(int count, int* first, int* n);

 private:
  std::mutex m_sampleMutex;
};

// PMRenderJob (PhotonMapping.h:33-121): the per-pixel statistics live on the
// GPU inside the job's photon-mapping state.
class PMRenderJob : public RenderJob {
 public:
  ~PMRenderJob();
  float alpha = 0.4f;
  HastySppm* sppm = nullptr;  // created by the first renderPhotonThread
  std::mutex sppmMutex;
};

class RenderThreadData {
 public:
  unsigned int seed;
  std::atomic<bool> stoppedFlag;
  explicit RenderThreadData(unsigned int s) : seed(s), stoppedFlag(false) {}
};

struct AccumulationBuffer {
  std::vector<float> data;  // W*H*3
  int numSamples = 0;
  uint32_t width = 0, height = 0;
  AccumulationBuffer(uint32_t w, uint32_t h) : data((size_t)w * h * 3, 0.0f), width(w), height(h) {}
};

template <typename Data>
class DoubleBuffer {
 public:
  DoubleBuffer(unsigned int width, unsigned int height) : m_write(width, height), m_read(width, height) {}
  void lock() { m_mutex.lock(); }
  void unlock() { m_mutex.unlock(); }
  Data& getWriteBuffer() { return m_write; }
  Data& getReadBuffer() { return m_read; }
  void copyBuffer() {
    std::lock_guard<std::mutex> lock(m_mutex);
    m_read = m_write;
  }

 private:
  std::mutex m_mutex;
  Data m_write;
  Data m_read;
};
typedef DoubleBuffer<AccumulationBuffer> Image3fAccDoubleBuffer;

// Samples are pulled from the job in batches of `batch` indices (0: all
// that remain) and rendered in one GPU call each into device films; after
// each batch a snapshot of the films is read back asynchronously (the next
// batch renders meanwhile) and, once it has arrived, published to the
// buffers (copyBuffer).  The write buffers hold sums and numSamples counts
// the samples in them, as if the reference's loop had run once per index.
// With `films` given, the thread renders into that film set (which the
// caller keeps, e.g. for hasty_films_estimate) instead of its own.
void renderThread(Image3fAccDoubleBuffer& colorBuffer, Image3fAccDoubleBuffer& normalBuffer,
                  Image3fAccDoubleBuffer& albedoBuffer, RenderJob& job, RenderThreadData& data, int batch = 0,
                  HastyFilms* films = nullptr);

// One photon-mapping pass per fetched sample index; the write buffers hold
// the current estimate with numSamples = 1.
void renderPhotonThread(Image3fAccDoubleBuffer& colorBuffer, Image3fAccDoubleBuffer& normalBuffer,
                        Image3fAccDoubleBuffer& albedoBuffer, PMRenderJob& job, RenderThreadData& threadData);

void computeEstimate(const std::vector<std::unique_ptr<Image3fAccDoubleBuffer>>& v, std::vector<float>& image);

// Colour PFM (bottom row first, as the format stores it).
void writePFM(const std::string& path, const std::vector<float>& rgb, uint32_t width, uint32_t height);

// writeEXR (core/src/Image.cpp:140-152) through hasty_image_write_exr.
void writeEXR(const std::string& path, const std::vector<float>& rgb, uint32_t width, uint32_t height);

}  // namespace Hasty

// C++ host layer: the reference's render-thread interface over the C ABI.
#include "hasty_host.h"

#include <stdio.h>

#include <algorithm>
#include <stdexcept>

namespace Hasty {

namespace {
void check(int rc) {
  if (rc != 0) throw std::runtime_error(hasty_last_error());
}
}  // namespace

GpuScene::GpuScene(const std::string& path) { check(hasty_scene_load(path.c_str(), &h_)); }
GpuScene::~GpuScene() {
  if (h_) hasty_scene_destroy(h_);
}

// RenderJob::loadJSON (RenderJob.cpp:21-28) + from_json (RenderJob.cpp:10-20)
void RenderJob::loadJSON(const std::string& path) {
  scene = std::make_shared<GpuScene>(path);
  HastySceneInfo info;
  check(hasty_scene_info(scene->handle(), &info));
  renderSettings.rouletteStartDepth = info.settings.roulette_start_depth;
  renderSettings.rouletteQ = info.settings.roulette_q;
  renderSettings.width = info.settings.width;
  renderSettings.height = info.settings.height;
  renderSettings.numThreads = info.settings.num_threads;
  renderSettings.maxDepth = info.settings.max_depth;
  renderSettings.numSamples = info.settings.num_samples;
  renderSettings.useMIS = info.settings.use_mis != 0;
  cameraExposure = info.camera_exposure;
}

void RenderJob::fetchSamples(int count, int* first, int* n) {
  std::lock_guard<std::mutex> lock(m_sampleMutex);
  int f = sampleIdxGenerator.load();
  int remaining = std::max(0, renderSettings.numSamples - f);
  int k = count > 0 ? std::min(count, remaining) : remaining;
  sampleIdxGenerator = f + k;
  *first = f;
  *n = k;
}

PMRenderJob::~PMRenderJob() {
  if (sppm) hasty_sppm_destroy(sppm);
}

// PathTracing.cpp:98-160 in batches of sample indices.
void renderThread(Image3fAccDoubleBuffer& colorBuffer, Image3fAccDoubleBuffer& normalBuffer,
                  Image3fAccDoubleBuffer& albedoBuffer, RenderJob& job, RenderThreadData& data, int batch,
                  HastyFilms* films) {
  Image3fAccDoubleBuffer* bufs[3] = {&colorBuffer, &normalBuffer, &albedoBuffer};
  for (auto* b : bufs) {
    std::fill(b->getWriteBuffer().data.begin(), b->getWriteBuffer().data.end(), 0.0f);
    b->getWriteBuffer().numSamples = 0;
  }
  HastyFilms* own = nullptr;
  if (!films) {
    std::lock_guard<std::mutex> lock(job.scene->mutex());
    if (hasty_films_create(job.scene->handle(), &own) != 0) {
      fprintf(stderr, "error: %s\n", hasty_last_error());
      data.stoppedFlag = true;
      return;
    }
    films = own;
  }
  // the newest finished snapshot, if any, into the write buffers -> readers
  auto publish = [&](int wait) {
    uint32_t k = 0;
    int rc = hasty_films_read(films, wait, colorBuffer.getWriteBuffer().data.data(),
                              normalBuffer.getWriteBuffer().data.data(), albedoBuffer.getWriteBuffer().data.data(), &k);
    if (rc == 0) {
      for (auto* b : bufs) {
        b->getWriteBuffer().numSamples = (int)k;
        b->copyBuffer();
      }
    } else if (rc != 4) {
      fprintf(stderr, "error: %s\n", hasty_last_error());
    }
  };
  while (!job.stopFlag) {
    int first, n;
    job.fetchSamples(batch, &first, &n);
    if (n <= 0) break;
    {
      std::lock_guard<std::mutex> lock(job.scene->mutex());
      int rc = hasty_films_render(films, data.seed, (uint32_t)first, (uint32_t)n);
      if (rc == 0) rc = hasty_films_snapshot(films);
      if (rc != 0) {
        fprintf(stderr, "error: %s\n", hasty_last_error());
        break;
      }
    }
    publish(0);
  }
  publish(1);  // the last snapshot, once its copy has arrived
  if (own) {
    std::lock_guard<std::mutex> lock(job.scene->mutex());
    hasty_films_destroy(own);
  }
  data.stoppedFlag = true;
}

// PhotonMapping.cpp:814-933
void renderPhotonThread(Image3fAccDoubleBuffer& colorBuffer, Image3fAccDoubleBuffer& normalBuffer,
                        Image3fAccDoubleBuffer& albedoBuffer, PMRenderJob& job, RenderThreadData& threadData) {
  Image3fAccDoubleBuffer* bufs[3] = {&colorBuffer, &normalBuffer, &albedoBuffer};
  for (auto* b : bufs) std::fill(b->getWriteBuffer().data.begin(), b->getWriteBuffer().data.end(), 0.0f);
  {
    std::lock_guard<std::mutex> lock(job.sppmMutex);
    if (!job.sppm) {
      HastySppmParams p{job.alpha, 10.0f, 0.5, 10.0, 0.0f};
      std::lock_guard<std::mutex> slock(job.scene->mutex());
      if (hasty_sppm_create(job.scene->handle(), threadData.seed, &p, &job.sppm) != 0) {
        fprintf(stderr, "error: %s\n", hasty_last_error());
        threadData.stoppedFlag = true;
        return;
      }
    }
  }
  RenderJobSample sample = job.fetchSample();
  while (!job.stopFlag && sample.idx < job.renderSettings.numSamples) {
    {
      std::lock_guard<std::mutex> lock(job.scene->mutex());
      int rc = hasty_sppm_pass(job.sppm, (uint32_t)sample.idx, colorBuffer.getWriteBuffer().data.data(),
                               normalBuffer.getWriteBuffer().data.data(), albedoBuffer.getWriteBuffer().data.data());
      if (rc != 0) {
        fprintf(stderr, "error: %s\n", hasty_last_error());
        break;
      }
    }
    for (auto* b : bufs) {
      b->getWriteBuffer().numSamples = 1;
      b->copyBuffer();
    }
    sample = job.fetchSample();
  }
  threadData.stoppedFlag = true;
}

// ViewerUtils.cpp:95-121
void computeEstimate(const std::vector<std::unique_ptr<Image3fAccDoubleBuffer>>& v, std::vector<float>& image) {
  int total = 0;
  for (size_t t = 0; t < v.size(); ++t) {
    auto& b = *v[t];
    b.lock();
    const auto& rb = b.getReadBuffer();
    if (total == 0) {  // as the reference: replaces until a buffer has samples
      image = rb.data;
    } else {
      for (size_t k = 0; k < image.size(); ++k) image[k] += rb.data[k];
    }
    total += rb.numSamples;
    b.unlock();
  }
  float d = 1e-7f + float(total);
  for (float& x : image) x /= d;
}

void writeEXR(const std::string& path, const std::vector<float>& rgb, uint32_t width, uint32_t height) {
  check(hasty_image_write_exr(path.c_str(), width, height, rgb.data()));
}

void writePFM(const std::string& path, const std::vector<float>& rgb, uint32_t width, uint32_t height) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  fprintf(f, "PF\n%u %u\n-1.0\n", width, height);
  for (uint32_t r = 0; r < height; ++r) {
    const float* row = rgb.data() + (size_t)(height - 1 - r) * width * 3;
    fwrite(row, sizeof(float), (size_t)width * 3, f);
  }
  fclose(f);
}

}  // namespace Hasty

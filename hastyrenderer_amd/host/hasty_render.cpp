// Headless renderer with the reference viewers' command line
// (apps/PathTracingViewer.cpp:20-166, apps/PhotonMappingViewer.cpp):
//   hasty_render [--sppm] <renderJobFile> <outputDirectory>
// starts numThreads render threads with seeds 0..numThreads-1, waits for
// them, and writes output.exr (colour estimate times 2^exposure),
// outputNormal.exr and outputAlbedo.exr as postProcessAndSaveToDisk does
// (apps/ViewerUtils.cpp:137-152; denoising stays outside this renderer).  The
// path tracer's estimate is computed on the GPU from the threads' device
// films (hasty_films_estimate).
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <memory>
#include <thread>
#include <vector>

#include "hasty_host.h"

using namespace Hasty;

int main(int argc, char** argv) {
  bool sppm = false;
  std::vector<const char*> args;
  for (int i = 1; i < argc; ++i) {
    if (strcmp(argv[i], "--sppm") == 0) sppm = true;
    else args.push_back(argv[i]);
  }
  if (args.size() != 2) {
    printf("unexpected number of arguments: %d expecting 3, usage <binary> [--sppm] <renderJobFile> <outputDirectory>\n",
           (int)args.size() + 1);
    return 1;
  }
  std::string out = args[1];
  struct stat st;
  if (stat(out.c_str(), &st) != 0 || !S_ISDIR(st.st_mode)) {
    printf("error given output path \"%s\" is not a directory\n", out.c_str());
    return 1;
  }
  std::unique_ptr<RenderJob> job = sppm ? std::unique_ptr<RenderJob>(new PMRenderJob()) : std::make_unique<RenderJob>();
  try {
    job->loadJSON(args[0]);
  } catch (const std::exception& e) {
    printf("exception: %s\n", e.what());
    return 2;
  }
  const RenderSettings& rs = job->renderSettings;
  unsigned nThreads = std::max(1u, rs.numThreads);
  std::vector<std::unique_ptr<RenderThreadData>> data;
  std::vector<std::unique_ptr<Image3fAccDoubleBuffer>> color, normal, albedo;
  std::vector<std::thread> threads;
  auto t0 = std::chrono::steady_clock::now();
  for (unsigned t = 0; t < nThreads; ++t) {
    data.emplace_back(std::make_unique<RenderThreadData>(t));
    color.emplace_back(std::make_unique<Image3fAccDoubleBuffer>(rs.width, rs.height));
    normal.emplace_back(std::make_unique<Image3fAccDoubleBuffer>(rs.width, rs.height));
    albedo.emplace_back(std::make_unique<Image3fAccDoubleBuffer>(rs.width, rs.height));
  }
  std::vector<HastyFilms*> films(sppm ? 0 : nThreads, nullptr);
  for (unsigned t = 0; t < films.size(); ++t) {
    if (hasty_films_create(job->scene->handle(), &films[t]) != 0) {
      printf("exception: %s\n", hasty_last_error());
      return 2;
    }
  }
  for (unsigned t = 0; t < nThreads; ++t) {
    if (sppm)
      threads.emplace_back(renderPhotonThread, std::ref(*color[t]), std::ref(*normal[t]), std::ref(*albedo[t]),
                           std::ref(*static_cast<PMRenderJob*>(job.get())), std::ref(*data[t]));
    else
      threads.emplace_back([&, t] { renderThread(*color[t], *normal[t], *albedo[t], *job, *data[t], 0, films[t]); });
  }
  for (auto& th : threads) th.join();
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const size_t P = (size_t)rs.width * rs.height * 3;
  std::vector<float> c(P), n(P), a(P);
  if (sppm) {
    computeEstimate(color, c);
    const float scale = std::pow(2.0f, job->cameraExposure);  // PathTracingViewer.cpp:160
    for (float& x : c) x *= scale;
    computeEstimate(normal, n);
    computeEstimate(albedo, a);
  } else {
    int rc = hasty_films_estimate(films.data(), (uint32_t)films.size(), job->cameraExposure, c.data(), n.data(),
                                  a.data());
    for (HastyFilms* f : films) hasty_films_destroy(f);
    if (rc != 0) {
      printf("exception: %s\n", hasty_last_error());
      return 5;
    }
  }
  try {
    writeEXR(out + "/output.exr", c, rs.width, rs.height);
    writeEXR(out + "/outputNormal.exr", n, rs.width, rs.height);
    writeEXR(out + "/outputAlbedo.exr", a, rs.width, rs.height);
  } catch (const std::exception& e) {
    printf("exception: %s\n", e.what());
    return 5;
  }
  printf("rendered %d %s of %ux%u in %.3f s\n", rs.numSamples, sppm ? "passes" : "samples", rs.width, rs.height,
         secs);
  return 0;
}

// ORACLE (test infrastructure only).  See oracle_sppm.h.
#include <climits>
#include "oracle_sppm.h"

#include <cmath>
#include <thread>

namespace oracle {

// HashCells.h:71-98: cells of the query box in i -> j -> k order, photons of a
// cell in insertion order, kept when |x - c|^2 <= r^2.
void PhotonMap::radiusNeighbors(const Vec3f& c, float r, std::vector<uint32_t>& out) const {
  out.clear();
  float r2 = r * r;
  CellKey lo = index(c - Vec3f::Fill(r));
  CellKey hi = index(c + Vec3f::Fill(r));
  for (int32_t i = lo.x; i <= hi.x; ++i)
    for (int32_t j = lo.y; j <= hi.y; ++j)
      for (int32_t k = lo.z; k <= hi.z; ++k) {
        auto it = cells.find(CellKey{i, j, k});
        if (it == cells.end()) continue;
        for (uint32_t idx : it->second)
          if (normSq(points[idx] - c) <= r2) out.push_back(idx);
      }
}

// Scene::sampleBRDFConcentrated (Scene.cpp:339-349) = sampleBRDF with the
// concentrated lobes only.
static SampleResult sampleConcentrated(RenderContext ctx, LightRayInfo& lightRay, const RayHit& hit, bool adjoint) {
  return ctx.scene.sampleBRDF(ctx.rng, lightRay, hit, adjoint, kConcentrated);
}

// PhotonMapping.cpp:191-265 (caustics map)
void tracePhoton(RenderContext ctx, PhotonMap& map, const Ray& a_ray, const Vec3f& a_flux) {
  Vec3f flux = a_flux;
  LightRayInfo lightRay;
  Ray ray = a_ray;
  RayHit hit;
  ctx.scene.rayHit(ray, hit);
  if (!hasHitSurface(hit)) return;
  Vec3f transmittance = beersLaw(lightRay.getTransmittance(), norm(hit.interaction.x - ray.origin()));
  flux = cwiseProd(flux, transmittance);
  for (int depth = 0; depth < ctx.settings.maxDepth; ++depth) {
    if (!isFinite(flux)) break;
    bool doStore = ctx.scene.hasBRDFDiffuseLobe(hit) && depth > 0;
    if (doStore) {
      Photon p;
      p.wi = -ray.direction();
      p.ngwi = dot(hit.interaction.normalGeometric, hit.wFrom());
      p.x = hit.interaction.x;
      p.flux = flux;
      map.photons.push_back(p);
      map.points.push_back(hit.interaction.x);
    }
    SampleResult s = sampleConcentrated(ctx, lightRay, hit, true);
    Vec3f throughputThis = s.throughput();
    if (throughputThis == Vec3f::Zero()) break;
    Ray ray2 = ctx.scene.constructRay(hit.interaction, s.direction, s.outside);
    RayHit hit2;
    ctx.scene.rayHit(ray2, hit2);
    if (!hasHitSurface(hit2) || isSelfIntersection(hit, hit2)) break;
    lightRay.updateMedia(ctx.scene, hit, s.direction);
    Vec3f t2 = beersLaw(lightRay.getTransmittance(), norm(hit2.interaction.x - ray2.origin()));
    throughputThis = cwiseProd(throughputThis, t2);
    float minQ = 1.0f - hm::powf_pos(0.1f, 1.0f / ctx.settings.maxDepth);
    float rouletteQ = std::max(minQ, 1.0f - norm(throughputThis));
    if (uniform01f(ctx.rng) <= rouletteQ) break;
    flux = cwiseProd(flux, throughputThis) / (1.0f - rouletteQ);
    hit = hit2;
    ray = ray2;
  }
}

// PhotonMapping.cpp:271-296 (radius > 0 on the active path)
static Vec3f photonMapFluxEstimate(RenderContext ctx, const PhotonMap& map, const RayHit& hit, float radius, int& M,
                                   std::vector<uint32_t>& found) {
  Vec3f est = Vec3f::Zero();
  map.radiusNeighbors(hit.interaction.x, radius, found);
  long long q[3] = {0, 0, 0};
  int flags = 0;  // bit 0: saturated, bit 1: non-finite term
  auto toFixed = [&](float t) -> long long {
    const float s = t * map.fluxScale;
    if (!(s == s) || std::isinf(t)) {
      flags |= 2;
      return 0;
    }
    if (!(std::fabs(s) < 0x1p62f)) {
      flags |= 1;
      return s < 0.0f ? -(1ll << 62) : (1ll << 62);
    }
    return std::llrint(s);  // round to nearest even
  };
  auto addSat = [&](long long a, long long b) -> long long {
    long long r;
    if (__builtin_add_overflow(a, b, &r)) {
      flags |= 1;
      return a < 0 ? LLONG_MIN : LLONG_MAX;
    }
    return r;
  };
  for (uint32_t idx : found) {
    const Photon& ph = map.photons[idx];
    MaterialEvalResult ev = ctx.scene.evaluteBRDF(hit, hit.wFrom(), ph.wi, 1.0f, false, kDiffuse);
    float fluxCorrection = std::fabs(dot(ph.wi, hit.interaction.normalGeometric)) / (1e-7f + std::fabs(ph.ngwi));
    Vec3f term = cwiseProd(ph.flux * fluxCorrection, ev.fDiffuse + ev.fConcentrated);
    if (map.fixedFlux) {
      for (int c = 0; c < 3; ++c) q[c] = addSat(q[c], toFixed(term[c]));
    } else {
      est += term;
    }
  }
  if (map.fixedFlux) {
    if (flags & 1) map.fluxOverflows.fetch_add(1);
    if (flags & 2) {
      est = Vec3f(NAN, NAN, NAN);
    } else {
      for (int c = 0; c < 3; ++c) est[c] = (float)((double)q[c] * map.fluxInvScale);
    }
  }
  M = (int)found.size();
  return est;
}

// PhotonMapping.cpp:404-503
static SppmResult pathTraceWithCausticsMap(RenderContext ctx, const PhotonMap& map, LightRayInfo& lightRay,
                                           const Ray& ray, RayHit* rayhitPtr, Vec3f& normal, Vec3f& albedo,
                                           float radius, int depth, std::vector<uint32_t>& found) {
  SppmResult result;
  if (depth >= ctx.settings.maxDepth) return result;
  RayHit local;
  RayHit& rayhit = rayhitPtr == nullptr ? local : *rayhitPtr;
  if (rayhitPtr == nullptr) ctx.scene.rayHit(ray, rayhit);
  if (hasHitSurface(rayhit)) {
    if (depth == 0) {
      normal = rayhit.interaction.normalShadingDefault;
      albedo = ctx.scene.getAlbedo(rayhit.interaction);
    }
    Vec3f throughput = beersLaw(lightRay.getTransmittance(), norm(rayhit.interaction.x - ray.origin()));
    if (ctx.scene.isSurfaceLight(rayhit)) {
      result.LLight += cwiseProd(throughput, ctx.scene.getEmissionRadiance(-ray.direction(), rayhit));
      return result;
    }
    if (ctx.scene.hasBRDFDiffuseLobe(rayhit)) {
      int M = 0;
      Vec3f est = photonMapFluxEstimate(ctx, map, rayhit, radius, M, found);
      result.M += M;
      result.photonFlux += cwiseProd(throughput, est);
    }
    bool twoStrategies = ctx.scene.hasLight() && ctx.settings.useMIS;
    int numStrategies = twoStrategies ? 2 : 1;
    int lightIndex = twoStrategies ? 0 : -1;
    float pdfs[2] = {0.0f, 0.0f};
    float rouletteQ = 0.0f;
    if (depth >= ctx.settings.rouletteStartDepth) rouletteQ = ctx.settings.rouletteQ;
    if (uniform01f(ctx.rng) <= rouletteQ) return result;
    throughput /= (1.0f - rouletteQ);
    for (int strategy = 0; strategy < numStrategies; ++strategy) {
      bool isLight = strategy == lightIndex;
      MISSampleResult s =
          isLight ? lightStrategySample(ctx, lightRay, ray, rayhit) : brdfStrategySample(ctx, lightRay, rayhit);
      Vec3f throughputWavelength = Vec3f::Ones();
      pdfs[strategy] = s.pdfOmega;
      if (s.throughput() == Vec3f::Zero() || isSelfIntersection(rayhit, s.rayhit)) continue;
      for (int j = 0; j < numStrategies; ++j) {
        if (j == strategy) continue;
        if (j == lightIndex) pdfs[j] = ctx.scene.evalLightRayFromStartPointDensity(rayhit.interaction, s.ray, s.rayhit);
        else pdfs[j] = ctx.scene.evaluateSamplePDF(rayhit, s.lightRay, s.ray.direction());
      }
      float msiWeight = msiWeightPowerHeuristic2(strategy, pdfs, numStrategies);
      s.lightRay.updateMedia(ctx.scene, rayhit, s.ray.direction());
      Vec3f tD = cwiseProd(throughput, cwiseProd(s.throughputDiffuse, throughputWavelength)) * msiWeight;
      Vec3f tC = cwiseProd(throughput, cwiseProd(s.throughputConcentrated, throughputWavelength)) * msiWeight;
      Vec3f t2 = tD + tC;
      Vec3f n2, a2;
      SppmResult rs =
          pathTraceWithCausticsMap(ctx, map, s.lightRay, s.ray, &s.rayhit, n2, a2, radius, depth + 1, found);
      result.LSpecular += cwiseProd(rs.LLight + rs.LSpecular, tC);
      result.LCaustic += cwiseProd(rs.LSpecular, tD);
      result.LRest += cwiseProd(rs.LRest, t2) + cwiseProd(rs.LLight, tD);
      result.photonFlux += cwiseProd(rs.photonFlux, t2);
      result.M += rs.M;
    }
    return result;
  }
  Vec3f transmittance = beersLaw(lightRay.getTransmittance(), 1e6f);
  Vec3f Le = ctx.scene.evalEnvironment(ray);
  result.LRest = cwiseProd(transmittance, Le);
  normal = Vec3f::Zero();
  albedo = result.LRest;
  return result;
}

// PhotonMapping.cpp:522-576
float computeCellSize(const std::vector<float>& radii, uint32_t W, uint32_t H) {
  float sum = 0.0f;
  for (uint32_t i = 0; i < W; ++i)
    for (uint32_t j = 0; j < H; ++j) sum += radii[(size_t)j * W + i];
  float a = 5.0f;
  float density = 1e5;
  float cellSize = sum / (W * H);
  auto computeCost = [&](float cs) {
    float s = 0.0f;
    for (uint32_t i = 0; i < W; ++i)
      for (uint32_t j = 0; j < H; ++j) {
        float r = radii[(size_t)j * W + i];
        s += a * powci<3>(r / cs) + powci<3>(std::max(r, cs)) * density;
      }
    return s;
  };
  float cost = computeCost(cellSize);
  for (int j = 0; j < 10; ++j) {
    float trial = cellSize * 1.5f;
    float trialCost = computeCost(trial);
    if (!(trialCost < cost)) break;
    cellSize = trial;
    cost = trialCost;
  }
  for (int j = 0; j < 10; ++j) {
    float trial = cellSize * 0.6666f;
    float trialCost = computeCost(trial);
    if (!(trialCost < cost)) break;
    cellSize = trial;
    cost = trialCost;
  }
  return cellSize;
}

// PhotonMapping.cpp:578-651 (ignoreCosineTerm = true)
float computeRadiusPixelFootPrint(RenderContext ctx, float defaultR, int i, int j, int W, int H) {
  Vec2f p0((float)i, (float)j);
  Vec2f offset(0.5f, 0.5f);
  Ray ray = ctx.scene.computeRay(ctx.rng, p0 + offset, float(W), float(H));
  RayHit rayhit;
  ctx.scene.rayHit(ray, rayhit);
  if (!hasHitSurface(rayhit)) return defaultR;
  float angle = ctx.scene.computeRayAngle(p0 + offset, float(W), float(H));
  const hasty::Camera& cam = ctx.scene.data.camera;
  Vec3f camPos(cam.position[0], cam.position[1], cam.position[2]);
  float d = norm(camPos - rayhit.interaction.x);
  if (!ctx.scene.hasBRDFDiffuseLobe(rayhit)) {
    for (int k = 0; k < ctx.settings.maxDepth; ++k) {
      if (ctx.scene.hasBRDFDiffuseLobe(rayhit)) break;
      SampleResult s;
      int tries = 0;
      do {
        LightRayInfo fresh;
        s = sampleConcentrated(ctx, fresh, rayhit, false);
        tries += 1;
      } while (s.throughputConcentrated == Vec3f::Zero() && tries < 3);
      if (s.throughputConcentrated == Vec3f::Zero()) break;
      Ray ray2 = ctx.scene.constructRay(rayhit.interaction, s.direction, s.outside);
      RayHit rayhit2;
      ctx.scene.rayHit(ray2, rayhit2);
      if (!hasHitSurface(rayhit2) || isSelfIntersection(rayhit, rayhit2)) return defaultR;
      d += norm(rayhit.interaction.x - rayhit2.interaction.x);
      rayhit = rayhit2;
    }
  }
  return 0.5f * d * hm::tanf(angle);
}

// PhotonMapping.cpp:653-678: rings of growing size; the fourth side of a
// ring (i2 = i - k + m, j2 = j + k) is computed but never tested.
float getCloseNonNegative(const std::vector<float>& img, int W, int H, int i, int j) {
  auto inside = [&](int x, int y) { return x >= 0 && y >= 0 && x < W && y < H; };
  auto at = [&](int x, int y) { return img[(size_t)y * W + x]; };
  int k = 1;
  while (k < std::max(H, W)) {
    for (int m = 0; m <= k * 2; ++m) {
      int i2 = i - k, j2 = j - k + m;
      if (inside(i2, j2) && at(i2, j2) > 0.0f) return at(i2, j2);
      i2 = i - k + m;
      j2 = j - k;
      if (inside(i2, j2) && at(i2, j2) > 0.0f) return at(i2, j2);
      i2 = i + k;
      j2 = j - k + m;
      if (inside(i2, j2) && at(i2, j2) > 0.0f) return at(i2, j2);
    }
    k += 1;
  }
  return 1e-3f;
}

// PhotonMapping.cpp:680-705
std::vector<float> computeInitialRadii(const PTJob& job, float footprintScale) {
  uint32_t W = job.settings.width, H = job.settings.height;
  std::vector<float> radii((size_t)W * H, -1.0f);
  int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto worker = [&](int t) {
    for (uint32_t p = (uint32_t)t; p < W * H; p += (uint32_t)threads) {
      uint32_t i = p % W, j = p / W;
      RNG rng = hrng::make_stream(job.seed, p, 0, hrng::kDomainRadius);
      RenderContext ctx{job.scene, rng, job.settings};
      radii[p] = footprintScale * computeRadiusPixelFootPrint(ctx, -1.0f, (int)i, (int)j, (int)W, (int)H);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(worker, t);
  for (auto& th : pool) th.join();
  for (uint32_t j = 0; j < H; ++j)
    for (uint32_t i = 0; i < W; ++i)
      if (radii[(size_t)j * W + i] < 0.0f) radii[(size_t)j * W + i] = getCloseNonNegative(radii, W, H, i, j);
  return radii;
}

// PMRenderJob::unlockAndUpdateStatistic (PhotonMapping.h:63-105); the pixel
// already holds its radius (lock, :45-54).
static Vec3f updateStatistic(PixelStat& px, float alpha, const Vec3f& flux, const Vec3f& L, int M, int NeInI,
                             float& outRadius) {
  px.LSum += L;
  px.Nd += 1;
  Vec3f Lflux = Vec3f::Zero();
  if (px.radius < 0.0f) {
    if (!(outRadius < 0.0f)) {
      px.radius = outRadius;
      px.N += alpha * M;
      px.Ne += NeInI;
      px.flux = flux;
      Lflux = px.Ne == 0 ? Vec3f::Zero() : px.flux / (px.radius * px.radius * float(Pi) * px.Ne);
    }
  } else {
    float rold = px.radius;
    float Ni = px.N;
    if (Ni + M > 0) px.radius *= hm::fsqrt((Ni + alpha * M) / (Ni + M));
    px.N += alpha * M;
    px.Ne += NeInI;
    px.flux = (px.flux + flux) * (px.radius * px.radius) / (rold * rold);
    Lflux = px.Ne == 0 ? Vec3f::Zero() : px.flux / (px.radius * px.radius * float(Pi) * px.Ne);
    outRadius = px.radius;
  }
  return Lflux + px.LSum / float(px.Nd);
}

void sppmPass(SppmJob& job, uint32_t pass, float* color, float* normal, float* albedo) {
  const PTJob& pt = job.pt;
  const Scene& scene = pt.scene;
  uint32_t W = pt.settings.width, H = pt.settings.height;
  if (!job.initialised) {
    job.pixels.assign((size_t)W * H, PixelStat());
    if (job.params.initialRadius > 0.0f) job.radii.assign((size_t)W * H, job.params.initialRadius);
    else if (scene.hasSurfaceLight()) job.radii = computeInitialRadii(pt, job.params.footprintScale);
    else job.radii.assign((size_t)W * H, 1.0f);
    job.initialised = true;
  }
  int emitLimit = (int)((float)job.params.emitFraction * H * W);
  int storeLimit = (int)((float)job.params.storeFraction * H * W);
  float cellSize = computeCellSize(job.radii, W, H);
  PhotonMap map;
  map.clear(cellSize);
  map.fixedFlux = job.params.fixedPointFlux;
  map.fluxScale = std::ldexp(1.0f, job.params.fluxFracBits);
  map.fluxInvScale = std::ldexp(1.0, -job.params.fluxFracBits);
  if (scene.hasSurfaceLight()) {
    for (int i = 0; i < emitLimit && (int)map.photons.size() < storeLimit; ++i) {
      RNG rng = hrng::make_stream(pt.seed, (uint32_t)i, pass, hrng::kDomainPhoton);
      RenderContext ctx{scene, rng, pt.settings};
      Vec3f flux;
      Ray ray = scene.sampleLightRay(rng, flux);
      map.emitted += 1;
      tracePhoton(ctx, map, ray, flux);
    }
  }
  map.build();
  job.lastEmitted = map.emitted;
  job.lastStored = map.photons.size();
  job.lastCellSize = cellSize;
  Vec2f offset = haltonOffset(pass);
  int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto worker = [&](int t) {
    std::vector<uint32_t> found;
    for (uint32_t p = (uint32_t)t; p < W * H; p += (uint32_t)threads) {
      uint32_t i = p % W, j = p / W;
      RNG rng = hrng::make_stream(pt.seed, p, pass, hrng::kDomainPath);
      RenderContext ctx{scene, rng, pt.settings};
      Ray ray = scene.computeRay(rng, Vec2f((float)i, (float)j) + offset, (float)W, (float)H);
      PixelStat& px = job.pixels[p];
      if (px.radius == -1.0f) px.radius = job.radii[p];
      float radius = px.radius;
      Vec3f n, a;
      LightRayInfo lightRay;
      SppmResult r = pathTraceWithCausticsMap(ctx, map, lightRay, ray, nullptr, n, a, radius, 0, found);
      Vec3f L = r.LLight + r.LRest + r.LSpecular;
      Vec3f Lnew;
      if (L != L || r.photonFlux != r.photonFlux) {
        Lnew = px.LSum / (1e-7f + float(px.Nd));
      } else {
        Lnew = updateStatistic(px, job.params.alpha, r.photonFlux, L, r.M, (int)map.emitted, radius);
      }
      job.radii[p] = radius;
      for (int c = 0; c < 3; ++c) {
        color[(size_t)p * 3 + c] = Lnew[c];
        normal[(size_t)p * 3 + c] = n[c];
        albedo[(size_t)p * 3 + c] = a[c];
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(worker, t);
  for (auto& th : pool) th.join();
  job.lastFluxOverflows = map.fluxOverflows.load();
  job.passesDone += 1;
}

}  // namespace oracle

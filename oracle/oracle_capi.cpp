// ORACLE (test infrastructure only).  C entry points used by tests/ (through
// ctypes), the smoke check and bench.py's CPU-baseline leg.  Nothing in the
// renderer links this library.
#include <string.h>

#include <memory>
#include <stdexcept>
#include <string>

#include "oracle_pt.h"
#include "oracle_sppm.h"

using namespace oracle;

static thread_local std::string g_err;

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// Loads renderJob.json (RenderJob::loadJSON semantics) and builds the BVH.
int oracle_job_load(const char* path, uint32_t seed, void** out) {
  try {
    auto job = std::make_unique<PTJob>();
    hasty::Job j;
    hasty::loadJob(path, j);
    job->scene.data = std::move(j.scene);
    job->scene.init();
    job->settings.rouletteStartDepth = j.settings.rouletteStartDepth;
    job->settings.rouletteQ = j.settings.rouletteQ;
    job->settings.width = j.settings.width;
    job->settings.height = j.settings.height;
    job->settings.maxDepth = j.settings.maxDepth;
    job->settings.numSamples = j.settings.numSamples;
    job->settings.useMIS = j.settings.useMIS;
    job->seed = seed;
    *out = job.release();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 2;
  }
}
void oracle_job_free(void* job) { delete (PTJob*)job; }

// settings overrides (width, height, max_depth, ...) for reduced-size checks
int oracle_job_set(void* jp, int width, int height, int maxDepth, int rrStart, float rrQ, int useMIS) {
  PTJob* job = (PTJob*)jp;
  if (width > 0) job->settings.width = (uint32_t)width;
  if (height > 0) job->settings.height = (uint32_t)height;
  if (maxDepth >= 0) job->settings.maxDepth = maxDepth;
  if (rrStart >= 0) job->settings.rouletteStartDepth = rrStart;
  if (rrQ >= 0.0f) job->settings.rouletteQ = rrQ;
  if (useMIS >= 0) job->settings.useMIS = useMIS != 0;
  return 0;
}
void oracle_job_info(void* jp, int* info) {
  PTJob* job = (PTJob*)jp;
  info[0] = (int)job->settings.width;
  info[1] = (int)job->settings.height;
  info[2] = job->settings.numSamples;
  info[3] = job->settings.maxDepth;
  info[4] = (int)job->scene.data.numTris;
  info[5] = (int)job->scene.data.lightTris.size();
  info[6] = job->scene.data.hasEnvTexture ? 1 : 0;
  info[7] = job->settings.rouletteStartDepth;
}

int oracle_render_pt(void* jp, uint32_t firstSample, uint32_t nSamples, int threads, float* color, float* normal,
                     float* albedo, uint64_t* nanCount) {
  try {
    renderPT(*(PTJob*)jp, firstSample, nSamples, threads, color, normal, albedo, nanCount, nullptr);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 3;
  }
}
// Same render with canonical BVH2 counters: out[0] rays, [1] node visits,
// [2] triangle tests.
int oracle_render_pt_counted(void* jp, uint32_t firstSample, uint32_t nSamples, float* color, float* normal,
                             float* albedo, uint64_t* out) {
  try {
    Bvh2::Stats st;
    uint64_t nans = 0;
    renderPT(*(PTJob*)jp, firstSample, nSamples, 1, color, normal, albedo, &nans, &st);
    out[0] = st.rays;
    out[1] = st.nodeVisits;
    out[2] = st.triTests;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 3;
  }
}

// Closest hit for n rays (origin xyz, direction xyz): tri (-1 miss), t, u, v,
// edge flag.  mode 0 = BVH2, 1 = brute force.
int oracle_trace(void* jp, const float* rays, uint32_t n, int mode, int32_t* tri, float* tuv, int32_t* edge) {
  PTJob* job = (PTJob*)jp;
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = rays + (size_t)i * 6;
    HitRecord h = mode == 0 ? job->scene.bvh.closest(r, r + 3, 1e-6f, nullptr) : job->scene.bvh.bruteForce(r, r + 3, 1e-6f);
    tri[i] = h.tri;
    tuv[i * 3 + 0] = h.t;
    tuv[i * 3 + 1] = h.u;
    tuv[i * 3 + 2] = h.v;
    edge[i] = h.edgeTie ? 1 : 0;
  }
  return 0;
}
// Material table entry: albedo rgb, roughness, metallic, specular, ior,
// anisotropy, transmission, emission rgb, varyIOR.
int oracle_material(void* jp, int idx, float* out) {
  PTJob* job = (PTJob*)jp;
  if (idx < 0 || idx >= (int)job->scene.data.materials.size()) return 1;
  const hasty::Material& m = job->scene.data.materials[idx];
  float v[14] = {m.albedo[0], m.albedo[1], m.albedo[2], m.roughness, m.metallic, m.specular, m.ior,
                 m.anisotropy, m.transmission, m.emission[0], m.emission[1], m.emission[2], m.varyIOR ? 1.0f : 0.0f,
                 0.0f};
  memcpy(out, v, sizeof(v));
  return 0;
}
// Texture indices of material idx: albedo, roughness, metallic, normal map.
int oracle_material_textures(void* jp, int idx, int32_t* out4) {
  PTJob* job = (PTJob*)jp;
  if (idx < 0 || idx >= (int)job->scene.data.materials.size()) return 1;
  const hasty::Material& m = job->scene.data.materials[idx];
  out4[0] = m.texAlbedo;
  out4[1] = m.texRoughness;
  out4[2] = m.texMetallic;
  out4[3] = m.texNormal;
  return 0;
}
// Texture t as loaded: width, height, channels; texels when px is not null.
int oracle_texture(void* jp, int t, uint32_t* whc, float* px) {
  PTJob* job = (PTJob*)jp;
  if (t < 0 || t >= (int)job->scene.data.textures.size()) return 1;
  const hasty::Texture& tx = job->scene.data.textures[t];
  whc[0] = tx.width;
  whc[1] = tx.height;
  whc[2] = tx.channels;
  if (px) memcpy(px, tx.px.data(), tx.px.size() * sizeof(float));
  return 0;
}
// Light table: global triangle index and area of light i.
int oracle_light(void* jp, int i, int32_t* tri, float* area) {
  PTJob* job = (PTJob*)jp;
  if (i < 0 || i >= (int)job->scene.data.lightTris.size()) return 1;
  *tri = (int32_t)job->scene.data.lightTris[i];
  *area = job->scene.data.lightAreas[i];
  return 0;
}
// Geometry ids of a global triangle index.
void oracle_tri_ids(void* jp, int32_t tri, uint32_t* geomPrim) {
  PTJob* job = (PTJob*)jp;
  geomPrim[0] = job->scene.data.triGeomID[tri];
  geomPrim[1] = job->scene.data.triPrimID[tri];
}

// Camera rays (Camera::computeRay) for pixel centres + Halton offset of sample k.
void oracle_camera_rays(void* jp, uint32_t sample, float* out) {
  PTJob* job = (PTJob*)jp;
  uint32_t W = job->settings.width, H = job->settings.height;
  Vec2f off = haltonOffset(sample);
  for (uint32_t j = 0; j < H; ++j)
    for (uint32_t i = 0; i < W; ++i) {
      RNG rng = hrng::make_stream(job->seed, j * W + i, sample, hrng::kDomainPath);
      Ray r = job->scene.computeRay(rng, Vec2f((float)i, (float)j) + off, (float)W, (float)H);
      float* o = out + ((size_t)j * W + i) * 6;
      for (int c = 0; c < 3; ++c) {
        o[c] = r.o[c];
        o[3 + c] = r.d[c];
      }
    }
}

// One camera ray through the continuous pixel position (px, py) with the
// lens stream (seed, 0, 0, 97): origin xyz, direction xyz.
void oracle_camera_ray(void* jp, float px, float py, uint32_t seed, float* out) {
  PTJob* job = (PTJob*)jp;
  RNG rng = hrng::make_stream(seed, 0, 0, 97);
  Ray r = job->scene.computeRay(rng, Vec2f(px, py), (float)job->settings.width, (float)job->settings.height);
  for (int c = 0; c < 3; ++c) {
    out[c] = r.o[c];
    out[3 + c] = r.d[c];
  }
}
// Environment radiance seen along n directions (Scene.cpp:482-485).
void oracle_env_eval(void* jp, const float* dirs, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* d = dirs + 3 * (size_t)i;
    Ray r(Vec3f(0.0f, 0.0f, 0.0f), Vec3f(d[0], d[1], d[2]));
    Vec3f L = ((PTJob*)jp)->scene.evalEnvironment(r);
    for (int c = 0; c < 3; ++c) out[3 * (size_t)i + c] = L[c];
  }
}
// Environment sampling density at n directions (Environment.h:158-165).
void oracle_env_pdf_n(void* jp, const float* dirs, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* d = dirs + 3 * (size_t)i;
    out[i] = ((PTJob*)jp)->scene.envEvalSamplingDensity(Vec3f(d[0], d[1], d[2]));
  }
}
// Positions of a loaded triangle (global index): p0, p1, p2.
int oracle_tri_pos(void* jp, uint32_t tri, float* out) {
  const PTJob* job = (const PTJob*)jp;
  if (tri >= job->scene.data.numTris) return 1;
  memcpy(out, &job->scene.data.triPos[(size_t)tri * 9], 9 * sizeof(float));
  return 0;
}
// Sets a planted mistake (see oracle::Mutation); 0 restores the restatement.
void oracle_set_mutation(int m) { g_mutation = m; }

// ---- leaf functions for unit tests -------------------------------------
void oracle_halton(uint32_t k, float* out) {
  Vec2f v = haltonOffset(k);
  out[0] = v[0];
  out[1] = v[1];
}
void oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  hrng::U32x4 r = hrng::philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  memcpy(out, r.v, 16);
}
// rows of (xi1, xi2, alphaX, alphaY, wOut xyz) -> Ne xyz
void oracle_sample_vndf(const float* in, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* a = in + i * 7;
    Vec3f r = sampleGGXVNDF(Vec3f(a[4], a[5], a[6]), a[2], a[3], a[0], a[1]);
    for (int c = 0; c < 3; ++c) out[i * 3 + c] = r[c];
  }
}
float oracle_dggx(float nh, float th, float bh, float at, float ab) { return DGGX(nh, th, bh, at, ab); }
float oracle_smith(float nv, float tv, float bv, float at, float ab) { return smithGGX(nv, tv, bv, at, ab); }
float oracle_fresnel(float ci, float ct, float ni, float nt) { return fresnelDielectric(ci, ct, ni, nt); }
int oracle_refract(const float* wi, const float* n, float ni, float nt, float* wt) {
  Vec3f o;
  bool ok = computeRefractionDirection(Vec3f(wi[0], wi[1], wi[2]), Vec3f(n[0], n[1], n[2]), ni, nt, o);
  for (int c = 0; c < 3; ++c) wt[c] = o[c];
  return ok ? 1 : 0;
}
// Draws n samples of a sampler from the stream (seed, 0, 0); writes dir xyz +
// pdf per sample and the pdf re-evaluated at that direction.
//   kind 0: cosine hemisphere about `n`     (param unused)
//   kind 1: GGX VNDF global, alpha = param, wo = w, frame from anyOrthonormal(n)
//   kind 2: uniform sphere
//   kind 3: n-gon aperture with int(param) blades (xy, pdf)
//   kind 4: triangle barycentrics (b0, b1, b2)
void oracle_sampler(int kind, uint32_t seed, uint32_t count, const float* nrm, const float* w, float param,
                    float* out) {
  RNG rng = hrng::make_stream(seed, 0, 0, 99);
  Vec3f n(nrm[0], nrm[1], nrm[2]);
  Vec3f wo(w[0], w[1], w[2]);
  for (uint32_t i = 0; i < count; ++i) {
    float* o = out + (size_t)i * 5;
    if (kind == 0) {
      float pdf;
      Vec3f d = sampleHemisphereCosImportance(rng, n, pdf);
      o[0] = d[0], o[1] = d[1], o[2] = d[2], o[3] = pdf, o[4] = evaluateHemisphereCosImportancePDF(n, d);
    } else if (kind == 1) {
      Vec3f t = anyOrthonormal(n);
      Vec3f b = cross(n, t);
      float pdf;
      Vec3f d = sampleGGXVNDFGlobal(rng, n, param, param, t, b, wo, pdf);
      o[0] = d[0], o[1] = d[1], o[2] = d[2], o[3] = pdf;
      o[4] = sampleGGXVNDFGlobalDensity(n, param, param, t, b, wo, d);
    } else if (kind == 2) {
      float pdf;
      Vec3f d = sampleSphereSurfaceUniformly(rng, pdf);
      o[0] = d[0], o[1] = d[1], o[2] = d[2], o[3] = pdf, o[4] = pdf;
    } else if (kind == 3) {
      float pdf;
      Vec2f p = sampleCircularNGonUniformly(rng, (int)param, 0.0f, pdf);
      o[0] = p[0], o[1] = p[1], o[2] = 0.0f, o[3] = pdf, o[4] = pdf;
    } else {
      Vec3f b = sampleTriangleUniformly(rng);
      o[0] = b[0], o[1] = b[1], o[2] = b[2], o[3] = 0.0f, o[4] = 0.0f;
    }
  }
}
// Environment sampling of a loaded job: n directions + sampling pdf + the
// evaluated pdf at that direction.
void oracle_env_sample(void* jp, uint32_t seed, uint32_t count, float* out) {
  PTJob* job = (PTJob*)jp;
  RNG rng = hrng::make_stream(seed, 0, 0, 98);
  for (uint32_t i = 0; i < count; ++i) {
    float pdf;
    Vec3f d = job->scene.envSample(rng, pdf);
    float* o = out + (size_t)i * 5;
    o[0] = d[0], o[1] = d[1], o[2] = d[2], o[3] = pdf, o[4] = job->scene.envEvalSamplingDensity(d);
  }
}
float oracle_env_pdf(void* jp, const float* dir) {
  return ((PTJob*)jp)->scene.envEvalSamplingDensity(Vec3f(dir[0], dir[1], dir[2]));
}
// Material from 9 parameters: albedo rgb, roughness, metallic, specular, ior,
// anisotropy, transmission.
static hasty::Material materialFromParams(const float* p) {
  hasty::Material m;
  m.albedo[0] = p[0], m.albedo[1] = p[1], m.albedo[2] = p[2];
  m.roughness = p[3];
  m.metallic = p[4];
  m.specular = p[5];
  m.ior = p[6];
  m.anisotropy = p[7];
  m.transmission = p[8];
  m.varyIOR = false;
  return m;
}
static SurfaceInteraction flatInteraction(const float* n) {
  SurfaceInteraction it;
  it.normalGeometric = Vec3f(n[0], n[1], n[2]);
  it.normalShadingDefault = it.normalGeometric;
  it.x = Vec3f(0.3f, 0.2f, 0.1f);
  it.tangent = orthonormalizedOtherwiseAnyOrthonormal(it.normalGeometric, Vec3f(-it.x[2], 0.0f, it.x[0]));
  it.bitangent = cross(it.normalGeometric, it.tangent);
  it.materialId = 0;
  return it;
}
// BSDF evaluate (both lobes) + sampling pdf of a parameterised material on a
// flat interaction with normal n at x = (0.3, 0.2, 0.1); rows of dirs6 are
// (wo, wi).  Output rows: fDiffuse xyz, fConcentrated xyz, pdf.
int oracle_bsdf_eval(const float* params, const float* n, const float* dirs6, uint32_t count, float* out7) {
  try {
    hasty::Material mat = materialFromParams(params);
    SurfaceInteraction it = flatInteraction(n);
    for (uint32_t i = 0; i < count; ++i) {
      const float* d = dirs6 + (size_t)i * 6;
      Vec3f vo(d[0], d[1], d[2]), vi(d[3], d[4], d[5]);
      MaterialEvalResult ev = Principled::evaluate(mat, it, vo, vi, 1.0f, false, kAll);
      float* o = out7 + (size_t)i * 7;
      for (int c = 0; c < 3; ++c) {
        o[c] = ev.fDiffuse[c];
        o[3 + c] = ev.fConcentrated[c];
      }
      o[6] = Principled::evaluateSamplePDF(mat, it, vo, vi, 1.0f);
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
// BSDF samples from the stream (seed, 0, 0, 97): direction xyz, throughput
// xyz (diffuse + concentrated), pdf.
int oracle_bsdf_sample(const float* params, const float* n, const float* wo, uint32_t seed, uint32_t count,
                       float* out7) {
  try {
    hasty::Material mat = materialFromParams(params);
    SurfaceInteraction it = flatInteraction(n);
    RNG rng = hrng::make_stream(seed, 0, 0, 97);
    for (uint32_t i = 0; i < count; ++i) {
      SampleResult s = Principled::sample(mat, rng, it, Vec3f(wo[0], wo[1], wo[2]), -1, 1.0f, false, kAll);
      float* o = out7 + (size_t)i * 7;
      Vec3f t = s.throughput();
      o[0] = s.direction[0], o[1] = s.direction[1], o[2] = s.direction[2];
      o[3] = t[0], o[4] = t[1], o[5] = t[2], o[6] = s.pdfOmega;
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
void oracle_rotation(const float* from, const float* to, const float* x, float* out) {
  RotationBetweenTwoVectors r(Vec3f(from[0], from[1], from[2]), Vec3f(to[0], to[1], to[2]));
  Vec3f y = applyRotation(r, Vec3f(x[0], x[1], x[2]));
  Vec3f z = applyRotationInverse(r, y);
  for (int c = 0; c < 3; ++c) {
    out[c] = y[c];
    out[3 + c] = z[c];
  }
}
void oracle_spherical(const float* dir, float* out) {
  Vec2f c = sphericalFromDirection(Vec3f(dir[0], dir[1], dir[2]));
  ToDirectionResult r = sphericalToDirection(c);
  out[0] = c[0], out[1] = c[1], out[2] = r.direction[0], out[3] = r.direction[1], out[4] = r.direction[2];
}
// ---- photon mapper ------------------------------------------------------
int oracle_sppm_create(const char* path, uint32_t seed, const float* params, void** out) {
  try {
    auto job = std::make_unique<SppmJob>();
    hasty::Job j;
    hasty::loadJob(path, j);
    job->pt.scene.data = std::move(j.scene);
    job->pt.scene.init();
    job->pt.settings.rouletteStartDepth = j.settings.rouletteStartDepth;
    job->pt.settings.rouletteQ = j.settings.rouletteQ;
    job->pt.settings.width = j.settings.width;
    job->pt.settings.height = j.settings.height;
    job->pt.settings.maxDepth = j.settings.maxDepth;
    job->pt.settings.numSamples = j.settings.numSamples;
    job->pt.settings.useMIS = j.settings.useMIS;
    job->pt.seed = seed;
    if (params) {  // alpha, footprint scale, store fraction, emit fraction, initial radius (<= 0: default)
      if (params[0] > 0) job->params.alpha = params[0];
      if (params[1] > 0) job->params.footprintScale = params[1];
      if (params[2] > 0) job->params.storeFraction = params[2];
      if (params[3] > 0) job->params.emitFraction = params[3];
      if (params[4] > 0) job->params.initialRadius = params[4];
      job->params.fixedPointFlux = params[5] != 0.0f;
      if (params[6] > 0) job->params.fluxFracBits = (int)params[6];
    }
    *out = job.release();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 2;
  }
}
void oracle_sppm_free(void* h) { delete (SppmJob*)h; }
int oracle_sppm_set_size(void* h, int width, int height) {
  SppmJob* job = (SppmJob*)h;
  if (job->initialised) return 1;
  job->pt.settings.width = (uint32_t)width;
  job->pt.settings.height = (uint32_t)height;
  return 0;
}
int oracle_sppm_pass(void* h, uint32_t pass, float* color, float* normal, float* albedo) {
  try {
    sppmPass(*(SppmJob*)h, pass, color, normal, albedo);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 3;
  }
}
// out: emitted, stored photons of the last pass; cell size; passes done
void oracle_sppm_info(void* h, double* out) {
  SppmJob* job = (SppmJob*)h;
  out[0] = (double)job->lastEmitted;
  out[1] = (double)job->lastStored;
  out[2] = job->lastCellSize;
  out[3] = job->passesDone;
  out[4] = job->lastFluxOverflows;
}
// per pixel: radius, N, Ne, flux rgb, LSum rgb, Nd (10 floats)
void oracle_sppm_pixels(void* h, float* out) {
  SppmJob* job = (SppmJob*)h;
  for (size_t p = 0; p < job->pixels.size(); ++p) {
    const PixelStat& px = job->pixels[p];
    float* o = out + p * 10;
    o[0] = px.radius, o[1] = px.N, o[2] = px.Ne;
    o[3] = px.flux[0], o[4] = px.flux[1], o[5] = px.flux[2];
    o[6] = px.LSum[0], o[7] = px.LSum[1], o[8] = px.LSum[2];
    o[9] = (float)px.Nd;
  }
}
// initial radii (computeInitialRadii) of the job's film
int oracle_sppm_initial_radii(void* h, float* out) {
  SppmJob* job = (SppmJob*)h;
  std::vector<float> r = computeInitialRadii(job->pt, job->params.footprintScale);
  memcpy(out, r.data(), sizeof(float) * r.size());
  return 0;
}
// the radius image the next pass starts from
void oracle_sppm_radii(void* h, float* out) {
  SppmJob* job = (SppmJob*)h;
  if (!job->radii.empty()) memcpy(out, job->radii.data(), sizeof(float) * job->radii.size());
}
float oracle_get_close(const float* img, int W, int H, int i, int j) {
  std::vector<float> v(img, img + (size_t)W * H);
  return getCloseNonNegative(v, W, H, i, j);
}
// radiusNeighbors over `n` points (3 floats each) with the given cell size:
// writes the indices in query order, returns how many (at most maxOut).
int64_t oracle_radius_neighbors(const float* pts, uint32_t n, float cellSize, const float* c, float r, uint32_t* out,
                                int64_t maxOut) {
  PhotonMap map;
  map.clear(cellSize);
  for (uint32_t k = 0; k < n; ++k) map.points.push_back(Vec3f(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]));
  map.build();
  std::vector<uint32_t> found;
  map.radiusNeighbors(Vec3f(c[0], c[1], c[2]), r, found);
  int64_t m = std::min<int64_t>((int64_t)found.size(), maxOut);
  for (int64_t k = 0; k < m; ++k) out[k] = found[k];
  return (int64_t)found.size();
}
float oracle_cell_size(const float* radii, uint32_t W, uint32_t H) {
  std::vector<float> r(radii, radii + (size_t)W * H);
  return computeCellSize(r, W, H);
}
// Photons stored by emission indices [first, first+count) of pass `pass`, in
// emission order: rows of x xyz, wi xyz, ngwi, flux rgb, emitter index.
// Returns the number of photons written (at most maxOut).
int64_t oracle_trace_photons(void* h, uint32_t pass, uint32_t first, uint32_t count, float* out, int64_t maxOut) {
  SppmJob* job = (SppmJob*)h;
  PhotonMap map;
  map.clear(1.0f);
  int64_t n = 0;
  for (uint32_t i = first; i < first + count; ++i) {
    size_t before = map.photons.size();
    RNG rng = hrng::make_stream(job->pt.seed, i, pass, hrng::kDomainPhoton);
    RenderContext ctx{job->pt.scene, rng, job->pt.settings};
    Vec3f flux;
    Ray ray = job->pt.scene.sampleLightRay(rng, flux);
    tracePhoton(ctx, map, ray, flux);
    for (size_t k = before; k < map.photons.size() && n < maxOut; ++k, ++n) {
      const Photon& p = map.photons[k];
      float* o = out + n * 11;
      o[0] = p.x[0], o[1] = p.x[1], o[2] = p.x[2];
      o[3] = p.wi[0], o[4] = p.wi[1], o[5] = p.wi[2];
      o[6] = p.ngwi;
      o[7] = p.flux[0], o[8] = p.flux[1], o[9] = p.flux[2];
      o[10] = (float)i;
    }
  }
  return n;
}

void oracle_hm_vec(int fn, const float* a, const float* b, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    float x = a[i], y = b[i], r;
    switch (fn) {
      case 0: r = hm::sinf(x); break;
      case 1: r = hm::cosf(x); break;
      case 2: r = hm::tanf(x); break;
      case 3: r = hm::atan2f(x, y); break;
      case 4: r = hm::acosf(x); break;
      case 5: r = hm::powf_pos(x, y); break;
      case 6: r = hm::fmodf(x, y); break;
      case 7: r = hm::atanf(x); break;
      case 8: r = hm::fsqrt(x); break;
      case 9: r = x / y; break;
      default: r = (float)((double)x * (double)y + 1.0); break;
    }
    out[i] = r;
  }
}
float oracle_hm(int fn, float a, float b) {
  switch (fn) {
    case 0: return hm::sinf(a);
    case 1: return hm::cosf(a);
    case 2: return hm::tanf(a);
    case 3: return hm::atan2f(a, b);
    case 4: return hm::acosf(a);
    case 5: return hm::powf_pos(a, b);
    case 6: return hm::fmodf(a, b);
    default: return hm::atanf(a);
  }
}

}  // extern "C"

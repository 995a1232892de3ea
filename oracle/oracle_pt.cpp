// ORACLE (test infrastructure only).  See oracle_pt.h.
#include "oracle_pt.h"

#include <cmath>
#include <stdexcept>
#include <thread>

namespace oracle {

int g_mutation = kMutNone;


Vec2f haltonOffset(uint32_t k) {
  HaltonSequence s1(2), s2(3);
  Vec2f r;
  for (uint32_t i = 0; i <= k; ++i) r = Vec2f(s1.next(), s2.next());
  return r;
}

// ============================================================ Sample.cpp ===
// Sample.cpp:5-11
Vec2f sampleDiskUniformly(RNG& rng, float& pDensity) {
  float alpha = uniform01f(rng) * float(2.0 * Pi);
  float r = hm::fsqrt(uniform01f(rng));
  pDensity = float(InvPi);
  return Vec2f(r * hm::cosf(alpha), r * hm::sinf(alpha));
}
// Sample.cpp:12-42
Vec2f sampleCircularNGonUniformly(RNG& rng, int numBlades, float bladeRotation, float& pDensity) {
  float anglePerSector = float(2.0 * Pi / numBlades);
  Vec2f normal = Vec2f(hm::cosf(anglePerSector * 0.5f), hm::sinf(anglePerSector * 0.5f));
  float d = -normal[0];
  Vec2f samplePosition;
  bool outside;
  do {
    samplePosition = sampleDiskUniformly(rng, pDensity);
    float r = norm(samplePosition);
    float angle = hm::atan2f(samplePosition[1], samplePosition[0]);
    angle -= bladeRotation;
    angle += float(2.0 * Pi) * (1.0f + FLT_EPSILON);
    angle = hm::fmodf(angle, anglePerSector);
    Vec2f normalizedPosition = Vec2f(r * hm::cosf(angle), r * hm::sinf(angle));
    outside = (dot(normal, normalizedPosition) + d) > 0.0f;
  } while (outside);
  float areaSector = 0.5 * hm::sinf(anglePerSector);
  float areaNGon = numBlades * areaSector;
  pDensity = 1.0 / areaNGon;
  return samplePosition;
}
// Sample.cpp:43-54
Vec3f sampleSphereSurfaceUniformly(RNG& rng, float& pDensity) {
  pDensity = 0.25f * float(InvPi);
  while (true) {
    float a = 2.0f * uniform01f(rng) - 1.0f;
    float b = 2.0f * uniform01f(rng) - 1.0f;
    float c = 2.0f * uniform01f(rng) - 1.0f;
    Vec3f r(a, b, c);
    float rsq = normSq(r);
    if (rsq <= 1.0 && rsq > 1e-6f) return r / hm::fsqrt(rsq);
  }
}
// Sample.cpp:57-73
Vec3f sampleHemisphereSurfaceUniformly(RNG& rng, const Vec3f& normal, float& pDensity) {
  pDensity = 0.5f * float(InvPi);
  while (true) {
    float a = 2.0f * uniform01f(rng) - 1.0f;
    float b = 2.0f * uniform01f(rng) - 1.0f;
    float c = 2.0f * uniform01f(rng) - 1.0f;
    Vec3f r(a, b, c);
    float rsq = normSq(r);
    if (rsq <= 1.0f && rsq > 1e-6f) {
      if (dot(r, normal) >= 0.0f) return r / hm::fsqrt(rsq);
    }
  }
}
// Sample.cpp:77-108
Vec3f sampleHemisphereCosImportance(RNG& rng, const Vec3f& normal, float& pDensity) {
  Vec3f dir;
  float costheta;
  do {
    float xi1 = uniform01f(rng);
    float xi2 = uniform01f(rng);
    float phi = 2.0f * float(Pi) * xi2;
    costheta = hm::fsqrt(xi1);
    float sintheta = hm::fsqrt(1.0f - xi1);
    float cosphi = hm::cosf(phi);
    float sinphi = hm::sinf(phi);
    Vec3f rv(sintheta * cosphi, sintheta * sinphi, costheta);
    RotationBetweenTwoVectors rotation(Vec3f(0.0f, 0.0f, 1.0f), normal);
    dir = applyRotation(rotation, rv);
    dir = normalize(dir);
    costheta = dot(dir, normal);
  } while (costheta <= 0.0f);
  pDensity = float(InvPi) * costheta;
  return dir;
}
// Sample.cpp:109-117
float evaluateHemisphereCosImportancePDF(const Vec3f& normal, const Vec3f& direction) {
  float cosTheta = dot(normal, direction);
  if (cosTheta < 0.0f) return 0.0f;
  return float(InvPi) * cosTheta;
}
// Sample.cpp:160-180
Vec3f sampleTriangleUniformly(RNG& rng) {
  float s = uniform01f(rng);
  float t = hm::fsqrt(uniform01f(rng));
  float beta1 = 1.0f - t;
  float beta3 = t * s;
  float beta2 = std::max(0.0f, 1.0f - beta1 - beta3);
  return Vec3f(beta1, beta2, beta3);
}

// ============================================================== BRDF.cpp ===
// BRDF.cpp:11-30
bool computeRefractionDirection(const Vec3f& wi, const Vec3f& normal, float iorI, float iorT, Vec3f& wt) {
  float relIOR = iorI / iorT;
  float cos_i = dot(wi, normal);
  float sin_iSq = std::max(0.f, 1.f - cos_i * cos_i);
  float sin_tSq = (relIOR * relIOR) * sin_iSq;
  if (sin_tSq >= 1.0f) return false;
  float cos_t = hm::fsqrt(1.0f - sin_tSq);
  wt = (-relIOR) * wi + (relIOR * cos_i - cos_t) * normal;
  wt = normalize(wt);
  return true;
}
// BRDF.cpp:31-49
static bool computeRefractionCosT(float cos_i, float iorI, float iorT, float& cos_t) {
  float relIOR = iorI / iorT;
  float sin_iSq = std::max(0.f, 1.f - cos_i * cos_i);
  float sin_tSq = (relIOR * relIOR) * sin_iSq;
  if (sin_tSq >= 1.0f) return false;
  cos_t = hm::fsqrt(1.0f - sin_tSq);
  return true;
}
// BRDF.cpp:52-61
static Vec3f computeRefractionDirectionFromAngles(const Vec3f& wi, const Vec3f& normal, float r, float cos_i,
                                                  float cos_t) {
  Vec3f wt = (-r) * wi + (r * cos_i - cos_t) * normal;
  return normalize(wt);
}
// BRDF.cpp:74-81
float fresnelDielectric(float cos_i, float cos_t, float ni, float nt) {
  float rp = (nt * cos_i - ni * cos_t) / (nt * cos_i + ni * cos_t);
  float ro = (ni * cos_i - nt * cos_t) / (ni * cos_i + nt * cos_t);
  return 0.5f * (rp * rp + ro * ro);
}
// BRDF.cpp:149-153
float DGGX(float nh, float th, float bh, float at, float ab) {
  return float(InvPi) / (at * ab * square(square(th / at) + square(bh / ab) + square(nh)));
}
// BRDF.cpp:155-159
float smithGGX(float nv, float tv, float bv, float at, float ab) {
  float lambda = 0.5f * (hm::fsqrt(1.0f + (square(at * tv) + square(ab * bv)) / square(nv)) - 1.0f);
  return 1.0f / (1.0f + lambda);
}
// BRDF.cpp:163-185
Vec3f sampleGGXVNDF(Vec3f wOut, float alphaX, float alphaY, float xi1, float xi2) {
  Vec3f Vh = normalize(Vec3f(alphaX * wOut[0], alphaY * wOut[1], wOut[2]));
  Vec3f X(1.0f, 0.0f, 0.0f), Z(0.0f, 0.0f, 1.0f);
  Vec3f T1 = cross(Vh, Z);
  if (normSq(T1) < 1e-12f) T1 = X;
  else T1 = normalize(T1);
  Vec3f T2 = cross(T1, Vh);
  float r = hm::fsqrt(xi1);
  float a = 1.0f / (1.0f + Vh[2]);
  float phi = float(Pi) * ((xi2 < a) ? (xi2 / a) : (1.0f + (xi2 - a) / (1.0f - a)));
  float t1 = r * hm::cosf(phi);
  float t2 = r * hm::sinf(phi) * (xi2 < a ? 1.0f : Vh[2]);
  float t3 = hm::fsqrt(std::max(0.0f, 1.0f - t1 * t1 - t2 * t2));
  Vec3f Nh = t1 * T1 + t2 * T2 + t3 * Vh;
  return normalize(Vec3f(alphaX * Nh[0], alphaY * Nh[1], std::max(0.0f, Nh[2])));
}
// BRDF.cpp:186-213.  The two uniforms passed as arguments at :199 are drawn
// xi1 first (C++ leaves the order unspecified).
Vec3f sampleGGXVNDFGlobal(RNG& rng, const Vec3f& normal, float at, float ab, const Vec3f& tangent,
                          const Vec3f& bitangent, const Vec3f& dir1, float& pDensity) {
  if (dot(dir1, normal) <= 0.0f) {
    pDensity = 0.0f;
    return -normal;
  }
  Vec3f NeGlobal;
  do {
    Vec3f Ve(dot(tangent, dir1), dot(bitangent, dir1), dot(normal, dir1));
    float xi1 = uniform01f(rng);
    float xi2 = uniform01f(rng);
    Vec3f Ne = sampleGGXVNDF(Ve, at, ab, xi1, xi2);
    NeGlobal = tangent * Ne[0] + bitangent * Ne[1] + normal * Ne[2];
    NeGlobal = normalize(NeGlobal);
  } while (dot(NeGlobal, dir1) <= 0.0f);
  Vec3f dir2 = reflectAcross(dir1, NeGlobal);
  dir2 = normalize(dir2);
  float G1 = smithGGX(dot(normal, dir1), dot(tangent, dir1), dot(bitangent, dir1), at, ab);
  float D = DGGX(dot(normal, NeGlobal), dot(tangent, NeGlobal), dot(bitangent, NeGlobal), at, ab);
  float pDensityH = G1 * dot(dir1, NeGlobal) * D / dot(dir1, normal);
  pDensity = pDensityH / (1e-7f + 4.0f * std::fabs(dot(dir1, NeGlobal)));
  return dir2;
}
// BRDF.cpp:214-226
float sampleGGXVNDFGlobalDensity(const Vec3f& normal, float at, float ab, const Vec3f& tangent, const Vec3f& bitangent,
                                 const Vec3f& dir1, const Vec3f& dir2) {
  if (dot(dir1, normal) <= 0.0f) return 0.0f;
  Vec3f NeGlobal = normalize(dir1 + dir2);
  float G1 = smithGGX(dot(normal, dir1), dot(tangent, dir1), dot(bitangent, dir1), at, ab);
  float D = DGGX(dot(normal, NeGlobal), dot(tangent, NeGlobal), dot(bitangent, NeGlobal), at, ab);
  float pDensityH = G1 * std::max(0.0f, dot(dir1, NeGlobal)) * D / dot(dir1, normal);
  return pDensityH / (1e-7f + 4.0f * std::fabs(dot(dir1, NeGlobal)));
}
static inline bool sideOutside(float dotNW) { return dotNW >= 0.0f; }   // BRDF.cpp:234-237
static inline float signSide(float dotNW) { return dotNW < 0.0f ? -1.0f : 1.0f; }  // BRDF.cpp:229-232
// BRDF.cpp:239-242
static inline float computeNormalScale(const Vec3f& wi, const Vec3f& ns, const Vec3f& ng) {
  return std::fabs(dot(wi, ns)) / (1e-7f + std::fabs(dot(wi, ng)));
}
// BRDF.cpp:253-263
float Principled::computeAlpha(float roughness) {
  float alpha = roughness * roughness;
  if (alpha < 1e-2f && alpha != 0.0f) alpha = 1e-2f;
  return alpha;
}
// BRDF.cpp:264-307
Principled::Prob Principled::computeProbability(float metallic, float specular, float transmission, float cos_o,
                                                float ior_o, float ior_refr_o, int flags) {
  Prob r;
  r.cosT = 0.0f;
  r.refractionPossible = computeRefractionCosT(cos_o, ior_o, ior_refr_o, r.cosT);
  if (r.refractionPossible) r.FDaccurate = fresnelDielectric(cos_o, r.cosT, ior_o, ior_refr_o);
  else r.FDaccurate = 1.0f;
  float mDiffuse = (1.0f - transmission) * (1.0f - metallic) * (1.0f - 0.5f * specular);
  float sTransmission = r.FDaccurate;
  float sMetallic = 1.0f;
  float mSpecular = mix(mix(0.5f * specular, sTransmission, transmission), sMetallic, metallic);
  float mRefractive = (1.0 - r.FDaccurate) * transmission * (1.0 - metallic);
  if (!(flags & kDiffuse)) mDiffuse = 0.0f;
  if (!(flags & kConcentrated)) {
    mSpecular = 0.0f;
    mRefractive = 0.0f;
  }
  float mSum = mDiffuse + mSpecular + mRefractive;
  r.noStrategy = mSum == 0.0f;
  if (r.noStrategy) {
    r.pDiffuse = r.pSpecular = r.pRefractive = 0.0f;
  } else {
    r.pDiffuse = mDiffuse / mSum;
    r.pSpecular = mSpecular / mSum;
    r.pRefractive = mRefractive / mSum;
  }
  return r;
}
// BRDF.cpp:308-316 + interaction tangent frame
static void computeAnisotropyParameters(const SurfaceInteraction& it, const Vec3f& ns, float alpha, float aniso,
                                        float& at, float& ab, Vec3f& t, Vec3f& b) {
  float aspect = hm::fsqrt(1.0f - 0.9f * aniso);
  at = alpha / aspect;
  ab = alpha * aspect;
  RotationBetweenTwoVectors rot(it.normalGeometric, ns);
  t = applyRotation(rot, it.tangent);
  b = applyRotation(rot, it.bitangent);
}
// BRDF.cpp:319-326
static Vec3f evaluateDiffuse(float dotNsWi, const Vec3f& albedo, float normalScale, float metallic, float specular,
                             float transmission) {
  if (dotNsWi < 0.0f) return Vec3f::Zero();
  return albedo * (float(InvPi) * normalScale * (1.0f - 0.5f * specular) * (1.0 - transmission) * (1.0 - metallic));
}
// BRDF.cpp:328-394
static Vec3f evaluateConcentratedWithoutTransmission(float dotNsWi, const Vec3f& albedo, float metallic,
                                                     float specular, float transmission, const Vec3f& wo,
                                                     const Vec3f& wi, float alpha, float at, float ab,
                                                     const Vec3f& tangent, const Vec3f& bitangent, const Vec3f& ns,
                                                     float normalScale, bool forceHitSpecular, const Vec3f& FD) {
  if (dotNsWi < 0.0f) return Vec3f::Zero();
  Vec3f h = normalize(wo + wi);
  float lh = std::fabs(dot(wi, h));
  float cosNih5 = powci<5>(1.0f - lh);
  float nh = dot(ns, h);
  float nv = std::fabs(dot(ns, wo));
  float nl = std::fabs(dot(ns, wi));
  float Go = smithGGX(nv, dot(tangent, wo), dot(bitangent, wo), at, ab);
  float Gi = smithGGX(nl, dot(tangent, wi), dot(bitangent, wi), at, ab);
  float G = Go * Gi;
  float D;
  if (alpha == 0.0f) D = forceHitSpecular ? 4.0f * nl : 0.0f;
  else D = DGGX(nh, dot(tangent, h), dot(bitangent, h), at, ab);
  float specularFromDielectric = 0.5f * specular;
  Vec3f specFromTransmission = Vec3f::Zero();
  Vec3f specularMetallic = albedo * (1.0f - cosNih5) + Vec3f::Ones() * cosNih5;
  if (transmission > 0.0f) {
    if (alpha != 0.0f) throw std::runtime_error("not implemented");
    specFromTransmission = forceHitSpecular ? FD : Vec3f::Zero();
  }
  return mix(mix(Vec3f::Fill(specularFromDielectric), specFromTransmission, transmission), specularMetallic,
             metallic) *
         (D * G * normalScale / std::fabs(1e-7f + 4.0f * nv * nl));
}
// BRDF.cpp:589-603
float Principled::getIndexOfRefraction(const hasty::Material& m, int wavelength) {
  if (!m.varyIOR) return m.ior;
  switch (wavelength) {
    case 0: return m.ior * (1.43f / 1.45f);
    case 1: return m.ior * (1.45f / 1.45f);
    case 2: return m.ior * (1.47f / 1.45f);
    default: return std::numeric_limits<float>::quiet_NaN();
  }
}
// ========================================================== Texture.h ===
// Texture::evaluateUV, Repeat addressing in u and v (Texture.h:119-140).
// Guards where the reference is undefined: a coordinate that is not finite or
// not below 2^31 in magnitude reads as 0; a wrapped coordinate that rounds to
// 1 reads the last column / row.
static size_t texelIndex(const hasty::Texture& t, const Vec2f& uvIn) {
  float uu = uvIn[0], vv = uvIn[1];
  if (!(std::fabs(uu) < 2147483648.0f)) uu = 0.0f;
  if (!(std::fabs(vv) < 2147483648.0f)) vv = 0.0f;
  float u = fmod1p1(uu);
  float oneMinusV = fmod1p1(g_mutation == kMutTexVNoFlip ? vv : 1.0f - vv);
  int col = static_cast<int>(u * t.width);
  int row = static_cast<int>(oneMinusV * t.height);
  col = std::min(std::max(col, 0), (int)t.width - 1);
  row = std::min(std::max(row, 0), (int)t.height - 1);
  return (size_t)row * t.width + (size_t)col;
}
// Texture<T>::evaluate (Texture.cpp:9-18) with the NoUV values of
// TextureTraits (Texture.h:84-101).
static Vec3f evaluateTexture3(const hasty::Texture& t, const SurfaceInteraction& it) {
  if (!it.hasUV) return Vec3f(1.0f, 0.0f, 0.7f);
  const float* p = &t.px[texelIndex(t, it.uv) * 3];
  return Vec3f(p[0], p[1], p[2]);
}
static float evaluateTexture1(const hasty::Texture& t, const SurfaceInteraction& it) {
  if (!it.hasUV) return -1.0f;
  return t.px[texelIndex(t, it.uv)];
}
// albedo / roughness / metallic ->evaluate(interaction) (BRDF.cpp:647-650,
// 685-687, 708-710, 745): a copy of the material with the textured values.
Scene::MatAt::MatAt(const Scene& scene, const SurfaceInteraction& it) {
  const hasty::Material& m = scene.mat(it);
  if (!m.textured()) {
    p = &m;
    return;
  }
  local = m;
  if (m.texAlbedo >= 0) {
    Vec3f a = evaluateTexture3(scene.data.textures[m.texAlbedo], it);
    for (int c = 0; c < 3; ++c) local.albedo[c] = a[c];
  }
  if (m.texRoughness >= 0) local.roughness = evaluateTexture1(scene.data.textures[m.texRoughness], it);
  if (m.texMetallic >= 0) local.metallic = evaluateTexture1(scene.data.textures[m.texMetallic], it);
  p = &local;
}

// BRDF.cpp:604-628
Vec3f Principled::getShadingNormal(const SurfaceInteraction& it, const Vec3f& wo, float dotNgWo) {
  float sideNgWo = signSide(dotNgWo);
  Vec3f ns;
  if (it.normalMap) {
    Vec3f color = evaluateTexture3(*it.normalMap, it);
    Vec3f texNormal = g_mutation == kMutNormalMapNoRemap ? color : color * 2.0f - Vec3f::Ones();
    ns = normalize(it.tangent * texNormal[0] + it.bitangent * texNormal[1] + it.normalGeometric * texNormal[2]);
  } else {
    ns = it.normalShadingDefault;
  }
  float sideNsWo = signSide(dot(ns, wo));
  if (sideNgWo != sideNsWo) ns = it.normalGeometric;
  ns *= sideNgWo;
  return ns;
}
// BRDF.cpp:630-679
MaterialEvalResult Principled::evaluate(const hasty::Material& m, const SurfaceInteraction& it, const Vec3f& wo,
                                        const Vec3f& wi, float iorOutside, bool adjoint, int flags) {
  (void)iorOutside;
  (void)adjoint;
  float dotNgWo = dot(it.normalGeometric, wo);
  Vec3f ns = getShadingNormal(it, wo, dotNgWo);
  float dotNsWi = dot(ns, wi);
  MaterialEvalResult r;
  r.normalShading = ns;
  Vec3f albedo(m.albedo[0], m.albedo[1], m.albedo[2]);
  float alpha = computeAlpha(m.roughness);
  float normalScale = computeNormalScale(wi, ns, it.normalGeometric);
  if (flags & kDiffuse)
    r.fDiffuse = evaluateDiffuse(dotNsWi, albedo, normalScale, m.metallic, m.specular, m.transmission);
  if (flags & kConcentrated) {
    Vec3f FD = Vec3f::Fill(std::numeric_limits<float>::quiet_NaN());
    float at, ab;
    Vec3f t, b;
    computeAnisotropyParameters(it, ns, alpha, m.anisotropy, at, ab, t, b);
    r.fConcentrated = evaluateConcentratedWithoutTransmission(dotNsWi, albedo, m.metallic, m.specular,
                                                              m.transmission, wo, wi, alpha, at, ab, t, b, ns,
                                                              normalScale, false, FD);
  }
  return r;
}
// BRDF.cpp:397-579 (static sample) with the member wrapper 681-700
SampleResult Principled::sample(const hasty::Material& m, RNG& rng, const SurfaceInteraction& it, const Vec3f& wo,
                                int wavelength, float iorOutside, bool adjoint, int flags) {
  Vec3f albedo(m.albedo[0], m.albedo[1], m.albedo[2]);
  float roughness = m.roughness, metallic = m.metallic, specular = m.specular, aniso = m.anisotropy,
        transmission = m.transmission;
  Vec3f ns = getShadingNormal(it, wo, dot(it.normalGeometric, wo));
  float iorMat = getIndexOfRefraction(m, wavelength);

  SampleResult result;
  float dotNgWo = dot(it.normalGeometric, wo);
  float dotNsWo = dot(ns, wo);
  float alpha = computeAlpha(roughness);
  result.outside = sideOutside(dotNgWo);
  float ior_o = result.outside ? iorOutside : iorMat;
  float ior_refr_o = result.outside ? iorMat : iorOutside;
  Prob pr = computeProbability(metallic, specular, transmission, dotNsWo, ior_o, ior_refr_o, flags);
  if (pr.noStrategy) {
    result.direction = Vec3f(1.0f, 0.0f, 0.0f);
    result.throughputDiffuse = Vec3f::Zero();
    result.throughputConcentrated = Vec3f::Zero();
    result.pdfOmega = 0.0f;
    return result;
  }
  Vec3f FD = Vec3f::Fill(pr.FDaccurate);
  Vec3f wavelengthMask = Vec3f::Ones();
  if (wavelength != -1) {
    wavelengthMask = Vec3f::Zero();
    wavelengthMask[wavelength] = 1.0f;
  }
  float at, ab;
  Vec3f tangent, bitangent;
  computeAnisotropyParameters(it, ns, alpha, aniso, at, ab, tangent, bitangent);
  MaterialEvalResult ev;
  ev.normalShading = ns;
  float xi = uniform01f(rng);
  const uint32_t DIFF = 1, SPEC = 2, REFR = 3;
  uint32_t strategy = xi < pr.pRefractive ? REFR : (xi < (pr.pRefractive + pr.pSpecular) ? SPEC : DIFF);
  if (strategy == SPEC) {
    if (alpha == 0.0f) {
      result.direction = reflectAcross(wo, ns);
      result.pdfOmega = 1e12f;
      const Vec3f& wi = result.direction;
      float dotNsWi = dot(ns, wi);
      float normalScale = computeNormalScale(wi, ns, it.normalGeometric);
      ev.fDiffuse = Vec3f::Zero();
      ev.fConcentrated = 1e12f / pr.pSpecular *
                         evaluateConcentratedWithoutTransmission(dotNsWi, albedo, metallic, specular, transmission, wo,
                                                                 result.direction, alpha, at, ab, tangent, bitangent,
                                                                 ns, normalScale, true, FD);
    } else {
      float pdfSpec;
      result.direction = sampleGGXVNDFGlobal(rng, ns, at, ab, tangent, bitangent, wo, pdfSpec);
      float pdfDiffuse = evaluateHemisphereCosImportancePDF(ns, result.direction);
      result.pdfOmega = pr.pSpecular * pdfSpec + pr.pDiffuse * pdfDiffuse;
      const Vec3f& wi = result.direction;
      float dotNsWi = dot(ns, wi);
      float normalScale = computeNormalScale(wi, ns, it.normalGeometric);
      ev.fDiffuse = evaluateDiffuse(dotNsWi, albedo, normalScale, metallic, specular, transmission);
      ev.fConcentrated = evaluateConcentratedWithoutTransmission(dotNsWi, albedo, metallic, specular, transmission, wo,
                                                                 result.direction, alpha, at, ab, tangent, bitangent,
                                                                 ns, normalScale, false, FD);
    }
  } else if (strategy == DIFF) {
    float pdfDiffuse;
    result.direction = sampleHemisphereCosImportance(rng, ns, pdfDiffuse);
    const Vec3f& wi = result.direction;
    float dotNsWi = dot(ns, wi);
    float normalScale = computeNormalScale(wi, ns, it.normalGeometric);
    if (alpha == 0.0f) {
      result.pdfOmega = pr.pDiffuse * pdfDiffuse;
      ev.fDiffuse = evaluateDiffuse(dotNsWi, albedo, normalScale, metallic, specular, transmission);
      ev.fConcentrated = Vec3f::Zero();
    } else {
      float pdfSpec = sampleGGXVNDFGlobalDensity(ns, at, ab, tangent, bitangent, wo, result.direction);
      result.pdfOmega = pr.pSpecular * pdfSpec + (1.0f - pr.pSpecular) * pdfDiffuse;
      ev.fDiffuse = evaluateDiffuse(dotNsWi, albedo, normalScale, metallic, specular, transmission);
      ev.fConcentrated = evaluateConcentratedWithoutTransmission(dotNsWi, albedo, metallic, specular, transmission, wo,
                                                                 result.direction, alpha, at, ab, tangent, bitangent,
                                                                 ns, normalScale, false, FD);
    }
  } else {
    result.outside = !result.outside;
    result.pdfOmega = 1e12f;
    ev.fDiffuse = Vec3f::Zero();
    if (!pr.refractionPossible) {
      ev.fConcentrated = Vec3f::Zero();
      result.direction = Vec3f(1.0f, 0.0f, 0.0f);
    } else {
      float r = ior_o / ior_refr_o;
      Vec3f wi = computeRefractionDirectionFromAngles(wo, ns, r, dotNsWo, pr.cosT);
      result.direction = wi;
      float scale = adjoint ? 1.0f : square(r);
      float fc = result.pdfOmega / pr.pRefractive * scale * (1.0f - pr.FDaccurate) /
                 std::fabs(dot(result.direction, it.normalGeometric));
      ev.fConcentrated = fc * wavelengthMask;
    }
  }
  float mm = std::fabs(dot(result.direction, it.normalGeometric)) / result.pdfOmega;
  result.throughputDiffuse = ev.fDiffuse * mm;
  result.throughputConcentrated = ev.fConcentrated * mm;
  return result;
}
// BRDF.cpp:702-742
float Principled::evaluateSamplePDF(const hasty::Material& m, const SurfaceInteraction& it, const Vec3f& wo,
                                    const Vec3f& wi, float outsideIOR) {
  float dotNgWo = dot(it.normalGeometric, wo);
  Vec3f ns = getShadingNormal(it, wo, dotNgWo);
  float dotNsWo = dot(ns, wo);
  float alpha = computeAlpha(m.roughness);
  float iorMat = m.ior;
  bool fromOutside = sideOutside(dotNgWo);
  float ior_o = fromOutside ? outsideIOR : iorMat;
  float ior_t = fromOutside ? iorMat : outsideIOR;
  Prob pr = computeProbability(m.metallic, m.specular, m.transmission, dotNsWo, ior_o, ior_t, kAll);
  if (alpha == 0.0f) {
    float pdfDiffuse = evaluateHemisphereCosImportancePDF(ns, wi);
    return pr.pDiffuse * pdfDiffuse;
  }
  float at, ab;
  Vec3f t, b;
  computeAnisotropyParameters(it, ns, alpha, m.anisotropy, at, ab, t, b);
  float pdfSpec = sampleGGXVNDFGlobalDensity(ns, at, ab, t, b, wo, wi);
  float pdfDiffuse = evaluateHemisphereCosImportancePDF(ns, wi);
  return pr.pSpecular * pdfSpec + pr.pDiffuse * pdfDiffuse;
}
// BRDF.cpp:743-750
bool Principled::hasDiffuseLobe(const hasty::Material& m) {
  return (1.0f - 0.5f * m.specular) * (1.0 - m.transmission) * (1.0 - m.metallic) > 0.0f;
}

// ======================================================= LightRayInfo.cpp ===
// LightRayInfo.cpp:8-38
void LightRayInfo::updateMedia(const Scene& scene, const RayHit& hit, const Vec3f& wTo) {
  float ndotFrom = dot(hit.interaction.normalGeometric, hit.wFrom());
  float ndotTo = dot(hit.interaction.normalGeometric, wTo);
  if (ndotFrom >= 0.0f && ndotTo < 0.0f) {
    Medium m;
    m.transmittance = scene.getAlbedo(hit.interaction);
    m.indexOfRefraction = scene.getIORInside(hit, wavelength);
    m.geomID = (int)hit.interaction.geomID;
    media.push_back(m);
    recompute();
  } else if (ndotFrom < 0.0f && ndotTo >= 0.0f) {
    int leave = (int)hit.interaction.geomID;
    int idx = -1;
    for (size_t i = 0; i < media.size(); ++i)
      if (media[i].geomID == leave) {
        idx = (int)i;
        break;
      }
    if (idx != -1) {
      media.erase(media.begin() + idx);
      recompute();
    }
  }
}
// LightRayInfo.cpp:39-50
void LightRayInfo::recompute() {
  for (size_t i = 0; i < media.size(); ++i) {
    if (media[i].geomID != -1) {
      transmittance = media[i].transmittance;
      return;
    }
  }
  transmittance = Vec3f::Ones();
}
// LightRayInfo.cpp:52-79 with sampleWeights = Ones (Scene.cpp:332)
int LightRayInfo::sampleOrGetWavelength(RNG& rng, bool sample, float& sampleProbability) {
  if (wavelength == -1 && sample) {
    Vec3f ones = Vec3f::Ones();
    Vec3f probabilities = ones / sum(ones);
    float xi = uniform01f(rng);
    if (xi < probabilities[0]) {
      wavelength = 0;
      sampleProbability = probabilities[0];
    } else if (xi < probabilities[0] + probabilities[1]) {
      wavelength = 1;
      sampleProbability = probabilities[1];
    } else {
      wavelength = 2;
      sampleProbability = probabilities[2];
    }
  } else {
    sampleProbability = 1.0f;
  }
  return wavelength;
}
// LightRayInfo.cpp:80-125
float LightRayInfo::getOutsideIOR(const RayHit& hit) const {
  if (dot(hit.wFrom(), hit.interaction.normalGeometric) < 0.0f) {
    int leave = (int)hit.interaction.geomID;
    int idx = -1;
    for (size_t i = 0; i < media.size(); ++i)
      if (media[i].geomID == leave) {
        idx = (int)i;
        break;
      }
    if (idx == -1) return 1.0f;
    if (media.size() == 1) return 1.0f;
    if (idx == 0) return media[1].indexOfRefraction;
    return media[0].indexOfRefraction;
  }
  if (media.empty()) return 1.0f;
  return media[0].indexOfRefraction;
}

// ============================================================= Scene.cpp ===
void Scene::init() {
  bvh.build(data.triPos.data(), data.numTris);
  for (size_t i = 0; i < data.materials.size(); ++i) {
    const hasty::Material& m = data.materials[i];
    // BRDF.cpp:386-390 throws when rough transmission is shaded; raised at
    // load, for a roughness texture if any texel or a face without uv
    // (roughness -1) gives alpha != 0
    if (m.transmission > 0.0f && m.texRoughness < 0 && Principled::computeAlpha(m.roughness) != 0.0f)
      throw std::runtime_error("not implemented");
    if (m.transmission > 0.0f && m.texRoughness >= 0) {
      for (float r : data.textures[m.texRoughness].px)
        if (Principled::computeAlpha(r) != 0.0f) throw std::runtime_error("not implemented");
      for (uint32_t t = 0; t < data.numTris; ++t)
        if (data.triMaterial[t] == (int32_t)i && !data.triHasUV[t]) throw std::runtime_error("not implemented");
    }
  }
}
// Scene.cpp:202-255
SurfaceInteraction Scene::getInteraction(int32_t tri, const Vec3f& bary, const Vec3f& geomNormal) const {
  SurfaceInteraction it;
  const float* p = &data.triPos[(size_t)tri * 9];
  Vec3f xv[3] = {Vec3f(p[0], p[1], p[2]), Vec3f(p[3], p[4], p[5]), Vec3f(p[6], p[7], p[8])};
  it.tri = tri;
  it.geomID = data.triGeomID[tri];
  it.primID = data.triPrimID[tri];
  it.barycentricCoordinates = bary;
  it.x = bary[0] * xv[0] + bary[1] * xv[1] + bary[2] * xv[2];
  it.normalGeometric = geomNormal;
  it.materialId = data.triMaterial[tri];
  if (data.triSmooth[tri]) {
    const float* q = &data.triNrm[(size_t)tri * 9];
    Vec3f nv[3] = {Vec3f(q[0], q[1], q[2]), Vec3f(q[3], q[4], q[5]), Vec3f(q[6], q[7], q[8])};
    it.normalShadingDefault = bary[0] * nv[0] + bary[1] * nv[1] + bary[2] * nv[2];
    it.normalShadingDefault = normalize(it.normalShadingDefault);
    it.h = 0.0f;
    for (int i = 0; i < 3; ++i) {
      float nvng = dot(it.normalGeometric, nv[i]);
      float hi = nvng == 0.0f ? 1e6 : (dot(xv[i], nv[i]) - dot(it.x, nv[i])) / nvng;
      it.h += it.barycentricCoordinates[i] * it.barycentricCoordinates[i] * hi;
    }
  } else {
    it.normalShadingDefault = it.normalGeometric;
    it.h = 0.0f;
  }
  if (data.triHasUV[tri]) {  // collectTriangleUV (tinyobj.cpp:73-88)
    const float* q = &data.triUV[(size_t)tri * 6];
    Vec2f uv[3] = {Vec2f(q[0], q[1]), Vec2f(q[2], q[3]), Vec2f(q[4], q[5])};
    it.hasUV = true;
    it.uv = uv[0] * bary[0] + uv[1] * bary[1] + uv[2] * bary[2];
  }
  const hasty::Material& m = data.materials[it.materialId];
  if (m.texNormal >= 0) it.normalMap = &data.textures[m.texNormal];
  it.tangent = Vec3f(-it.x[2], 0.0f, it.x[0]);
  it.tangent = orthonormalizedOtherwiseAnyOrthonormal(it.normalGeometric, it.tangent);
  it.bitangent = cross(it.normalGeometric, it.tangent);
  return it;
}
// Scene.cpp:257-270
Vec3f Scene::constructRayX(const SurfaceInteraction& it, bool outside) const {
  Vec3f normal = it.normalGeometric;
  float eps = 1e-5f;
  float hmin = std::max(eps, norm(it.x) * eps);
  return it.x + normal * (1.01f * (outside ? std::max(0.0f, it.h) + hmin : std::min(0.0f, it.h) - hmin));
}
Ray Scene::constructRay(const SurfaceInteraction& it, const Vec3f& dir, bool outside) const {  // :271-275
  return Ray(constructRayX(it, outside), dir);
}
Ray Scene::constructRayEnd(const SurfaceInteraction& it, const Vec3f& end, bool outside) const {  // :276-280
  Vec3f xo = constructRayX(it, outside);
  return Ray(xo, normalize(end - xo));
}
// Scene.cpp:281-303
void Scene::rayHit(const Ray& ray, RayHit& hit) const {
  hit.rayDir = ray.direction();
  hit.rtc = bvh.closest(ray.origin().v, ray.direction().v, 1e-6f, stats);
  if (hasHitSurface(hit)) {
    Vec3f bary(1.0f - hit.rtc.u - hit.rtc.v, hit.rtc.u, hit.rtc.v);
    const float* p = &data.triPos[(size_t)hit.rtc.tri * 9];
    Vec3f p0(p[0], p[1], p[2]), p1(p[3], p[4], p[5]), p2(p[6], p[7], p[8]);
    Vec3f ng = normalize(cross(p1 - p0, p2 - p0));
    hit.interaction = getInteraction(hit.rtc.tri, bary, ng);
  }
}
Vec3f Scene::getAlbedo(const SurfaceInteraction& it) const {  // Scene.cpp:478-481
  MatAt m(*this, it);
  return Vec3f(m->albedo[0], m->albedo[1], m->albedo[2]);
}
bool Scene::hasBRDFDiffuseLobe(const RayHit& hit) const {
  return Principled::hasDiffuseLobe(*MatAt(*this, hit.interaction));
}
float Scene::getIORInside(const RayHit& hit, int wavelength) const {
  return Principled::getIndexOfRefraction(mat(hit.interaction), wavelength);
}
MaterialEvalResult Scene::evaluteBRDF(const RayHit& hit, const Vec3f& wo, const Vec3f& wi, float iorOutside,
                                      bool adjoint, int flags) const {  // :323-326
  return Principled::evaluate(*MatAt(*this, hit.interaction), hit.interaction, wo, wi, iorOutside, adjoint, flags);
}
// Scene.cpp:328-349
SampleResult Scene::sampleBRDF(RNG& rng, LightRayInfo& lightRay, const RayHit& hit, bool adjoint, int flags) const {
  MatAt matAt(*this, hit.interaction);
  const hasty::Material& m = *matAt;
  float sampleProbability;
  int wavelength = lightRay.sampleOrGetWavelength(rng, m.varyIOR, sampleProbability);
  float outsideIOR = lightRay.getOutsideIOR(hit);
  SampleResult r = Principled::sample(m, rng, hit.interaction, hit.wFrom(), wavelength, outsideIOR, adjoint, flags);
  r.throughputDiffuse /= sampleProbability;
  r.throughputConcentrated /= sampleProbability;
  return r;
}
// Scene.cpp:350-353
float Scene::evaluateSamplePDF(const RayHit& hit, const LightRayInfo& lightRay, const Vec3f& dir2) const {
  return Principled::evaluateSamplePDF(*MatAt(*this, hit.interaction), hit.interaction, hit.wFrom(), dir2,
                                       lightRay.getOutsideIOR(hit));
}
// Scene.cpp:494-500
bool Scene::isSurfaceLight(const RayHit& hit) const {
  if (!hasHitSurface(hit)) return false;
  return mat(hit.interaction).isLight();
}
// Scene.cpp:460-477
Vec3f Scene::getEmissionRadianceOfTri(int32_t tri) const {
  const hasty::Material& m = data.materials[data.triMaterial[tri]];
  return Vec3f(m.emission[0], m.emission[1], m.emission[2]) * float(InvPi);
}
Vec3f Scene::getEmissionRadiance(const Vec3f& wo, const RayHit& hit) const {
  if (dot(hit.interaction.normalGeometric, wo) <= 0.0f) return Vec3f::Zero();
  const hasty::Material& m = mat(hit.interaction);
  return Vec3f(m.emission[0], m.emission[1], m.emission[2]) * float(InvPi);
}
// Scene.cpp:482-485
Vec3f Scene::evalEnvironment(const Ray& ray) const {
  if (!data.hasEnvTexture) return Vec3f(data.backgroundColor[0], data.backgroundColor[1], data.backgroundColor[2]);
  Vec2f c = sphericalFromDirection(ray.direction());
  Vec2f uv(c[1] * float(0.5 * InvPi), 1.0f - c[0] * float(InvPi));
  return envEvaluateUV(uv);
}
// Texture.h:119-140 with U repeat, V clamp (Environment.h:115-116).  The
// column is clamped to width-1 where the reference would read past the row.
Vec3f Scene::envEvaluateUV(const Vec2f& uv) const {
  // A direction slightly off the unit sphere gives acos(NaN); the reference
  // then converts NaN to an index (undefined).  Zero radiance instead.
  if (uv[0] != uv[0] || uv[1] != uv[1]) return Vec3f::Zero();
  float u = fmod1p1(uv[0]);
  float oneMinusV = std::min(std::max(1.0f - uv[1], 0.0f), 1.0f - FLT_EPSILON);
  int col = (int)(u * data.env.width);
  int row = (int)(oneMinusV * data.env.height);
  if (col > (int)data.env.width - 1) col = (int)data.env.width - 1;
  if (g_mutation == kMutEnvRowFromBottom) row = (int)data.env.height - 1 - row;
  const float* px = &data.env.rgb[((size_t)row * data.env.width + col) * 3];
  return Vec3f(px[0], px[1], px[2]);
}
// Environment.h:59-78
Vec2f Scene::envSampleUV(RNG& rng, float& density) const {
  uint32_t n = (uint32_t)data.envCdf.size();
  uint32_t idx = hrng::upper_bound(data.envCdf.data(), n, rng.canonical_double());
  if (idx >= n) idx = n - 1;
  uint32_t W = data.env.width, H = data.env.height;
  uint32_t x = idx % W;
  uint32_t y = idx / W;
  float heightF = float(H);
  float widthF = float(W);
  float xi0 = uniform01f(rng);
  float xi1 = uniform01f(rng);
  float left = x / widthF;
  float right = (x + 1.0f) / widthF;
  float lower = (H - (y + 1.0f)) / heightF;
  float upper = (float)(H - y) / heightF;
  float u = (1.0f - xi0) * left + xi0 * right;
  float v = (1.0f - xi1) * lower + xi1 * upper;
  float pxy = data.envPdf[idx] / data.envValueSum;
  density = widthF * heightF * pxy;
  return Vec2f(u, v);
}
// Environment.h:80-87
float Scene::envEvalDensityUV(const Vec2f& uv) const {
  if (uv[0] != uv[0] || uv[1] != uv[1]) return 0.0f;  // see envEvaluateUV
  uint32_t W = data.env.width, H = data.env.height;
  float heightF = float(H), widthF = float(W);
  uint32_t x = (uint32_t)std::min((size_t)(uv[0] * W), (size_t)W - 1);
  uint32_t y = (uint32_t)std::min((size_t)((1.0f - uv[1]) * H), (size_t)H - 1);
  return widthF * heightF * data.envPdf[(size_t)y * W + x] / data.envValueSum;
}
// Environment.h:141-153
Vec3f Scene::envSample(RNG& rng, float& density) const {
  ToDirectionResult sc;
  do {
    Vec2f uv = envSampleUV(rng, density);
    Vec2f c;
    c[1] = uv[0] * 2.0f * float(Pi);
    c[0] = (1.0f - uv[1]) * float(Pi);
    sc = sphericalToDirection(c);
  } while (sc.sinTheta == 0.0f);
  density /= (2.0f * float(Pi) * float(Pi) * sc.sinTheta);
  return sc.direction;
}
// Environment.h:158-165
float Scene::envEvalSamplingDensity(const Vec3f& dir) const {
  Vec2f c = sphericalFromDirection(dir);
  Vec2f uv(c[1] * float(0.5 * InvPi), 1.0f - c[0] * float(InvPi));
  float sinTheta = hm::sinf(c[0]);
  if (g_mutation == kMutEnvPdfNoSinTheta) sinTheta = 1.0f;
  return envEvalDensityUV(uv) / std::max(1e-7f, 2.0f * float(Pi * Pi) * sinTheta);
}
// Scene.cpp:506-526
SurfaceInteraction Scene::sampleSurfaceLightPosition(RNG& rng, float& pDensity) const {
  if (data.lightTris.empty()) throw std::runtime_error("error: no light triangle found ");
  pDensity = 1.0f / data.totalLightArea;
  uint32_t n = (uint32_t)data.lightCdf.size();
  uint32_t idx = hrng::upper_bound(data.lightCdf.data(), n, rng.canonical_double());
  if (idx >= n) idx = n - 1;
  int32_t tri = (int32_t)data.lightTris[idx];
  const float* p = &data.triPos[(size_t)tri * 9];
  Vec3f p0(p[0], p[1], p[2]), p1(p[3], p[4], p[5]), p2(p[6], p[7], p[8]);
  Vec3f normal = normalize(cross(p1 - p0, p2 - p0));
  Vec3f bary = sampleTriangleUniformly(rng);
  return getInteraction(tri, bary, normal);
}
// Scene.cpp:527-537
Ray Scene::sampleLightRay(RNG& rng, Vec3f& flux) const {
  float pPos, pDirection;
  SurfaceInteraction it = sampleSurfaceLightPosition(rng, pPos);
  Vec3f direction = sampleHemisphereCosImportance(rng, it.normalGeometric, pDirection);
  Vec3f L = getEmissionRadianceOfTri(it.tri);
  flux = L * (std::fabs(dot(it.normalGeometric, direction)) / (pPos * pDirection));
  return constructRay(it, direction, true);
}
// Scene.cpp:538-586
Ray Scene::sampleLightRayFromStartPoint(RNG& rng, const SurfaceInteraction& point, float& pDensity, RayHit& hit,
                                        bool& lightVisible) const {
  Ray resultray;
  float pSurfaceSampling = !data.hasEnvTexture ? 1.0f : (hasSurfaceLight() ? 0.5f : 0.0f);
  float xi = uniform01f(rng);
  bool useSurfaceSampling = xi < pSurfaceSampling;
  if (useSurfaceSampling) {
    float pDensityA;
    SurfaceInteraction it2 = sampleSurfaceLightPosition(rng, pDensityA);
    Vec3f diff = it2.x - point.x;
    resultray = constructRayEnd(point, it2.x, dot(point.normalGeometric, diff) >= 0.0f);
    rayHit(resultray, hit);
    lightVisible = hasHitSurface(hit, it2.tri);
    if (lightVisible) pDensity = evalLightRayFromStartPointDensity(point, resultray, hit);
  } else {
    Vec3f direction = envSample(rng, pDensity);
    resultray = constructRay(point, direction, dot(point.normalGeometric, direction) >= 0.0f);
    rayHit(resultray, hit);
    lightVisible = isSurfaceLight(hit) || isInfiniteAreaLight(hit);
    if (lightVisible) pDensity = evalLightRayFromStartPointDensity(point, resultray, hit);
  }
  return resultray;
}
// Scene.cpp:587-627
float Scene::evalLightRayFromStartPointDensity(const SurfaceInteraction& point, const Ray& ray2,
                                               const RayHit& hit2) const {
  (void)point;
  bool surfaceLight = isSurfaceLight(hit2);
  bool infiniteAreaLight = isInfiniteAreaLight(hit2);
  if (!surfaceLight && !infiniteAreaLight) return 0.0f;
  float pSurfaceSampling = !data.hasEnvTexture ? 1.0f : (hasSurfaceLight() ? 0.5f : 0.0f);
  float resultDensity;
  if (surfaceLight) {
    float lightStrategyPdfA = 1.0f / data.totalLightArea;
    float diffSq = std::max(1e-5f, normSq(ray2.origin() - hit2.interaction.x));
    Vec3f normal2 = hit2.interaction.normalGeometric;
    float lightStrategyPdfOmega = lightStrategyPdfA * diffSq / (1e-7f + std::fabs(dot(normal2, ray2.direction())));
    resultDensity = pSurfaceSampling * lightStrategyPdfOmega;
    if (data.hasEnvTexture) resultDensity += (1.0f - pSurfaceSampling) * envEvalSamplingDensity(ray2.direction());
  } else {
    if (data.hasEnvTexture) resultDensity = (1.0f - pSurfaceSampling) * envEvalSamplingDensity(ray2.direction());
    else resultDensity = 0.0f;
  }
  return resultDensity;
}
// Camera.cpp:41-81
Ray Scene::computeRay(RNG& rng, const Vec2f& p, float frameWidth, float frameHeight) const {
  const hasty::Camera& c = data.camera;
  Vec3f pos(c.position[0], c.position[1], c.position[2]);
  Vec3f fwd(c.forward[0], c.forward[1], c.forward[2]);
  Vec3f right(c.right[0], c.right[1], c.right[2]);
  Vec3f up(c.up[0], c.up[1], c.up[2]);
  float x = p[0];
  float y = p[1];
  float fovRadians = float(Pi) * c.fovDegree / 180.0f;
  float q = hm::tanf(fovRadians * 0.5f);
  float lx = (x - 0.5f * frameWidth) / frameWidth * 2.0f;
  float ly = (y - 0.5f * frameHeight) / frameWidth * 2.0f;
  if (g_mutation == kMutCameraLyByHeight) ly = (y - 0.5f * frameHeight) / frameHeight * 2.0f;
  Vec3f direction = fwd + right * (q * lx) - up * (q * ly);
  if (g_mutation == kMutCameraUpSign) direction = fwd + right * (q * lx) + up * (q * ly);
  direction = normalize(direction);
  if (c.aperture > 0.0f) {
    float density;
    Vec2f offset = sampleCircularNGonUniformly(rng, c.numBlades, c.bladeRotation, density);
    float t = c.focalDistance * dot(direction, fwd);
    if (g_mutation == kMutLensFocalAlongRay) t = c.focalDistance;
    Vec3f focalPoint = pos + t * direction;
    Vec3f origin = pos + c.aperture * offset[0] * up + c.aperture * offset[1] * right;
    direction = normalize(focalPoint - origin);
    return Ray(origin, direction);
  }
  return Ray(pos, direction);
}
// Camera.cpp:83-98
float Scene::computeRayAngle(const Vec2f& p, float frameWidth, float frameHeight) const {
  const hasty::Camera& c = data.camera;
  float x = p[0];
  float y = p[1];
  float fovRadians = float(Pi) * c.fovDegree / 180.0f;
  float q = hm::tanf(fovRadians * 0.5f);
  float lx1 = (x + 0.5f - 0.5f * frameWidth) / frameWidth * 2.0f;
  float ly1 = (y + 0.5f - 0.5f * frameHeight) / frameWidth * 2.0f;
  float lx2 = (x - 0.5f - 0.5f * frameWidth) / frameWidth * 2.0f;
  float ly2 = (y - 0.5f - 0.5f * frameHeight) / frameWidth * 2.0f;
  float angleX = std::fabs(hm::atanf(q * lx1) - hm::atanf(q * lx2));
  float angleY = std::fabs(hm::atanf(q * ly1) - hm::atanf(q * ly2));
  return 0.5f * angleX + 0.5f * angleY;
}

// ======================================================= PathTracing.cpp ===
// PathTracing.h:46-59
Vec3f beersLaw(const Vec3f& transmittance, float d) {
  if (transmittance == Vec3f::Ones()) return Vec3f::Ones();
  if (g_mutation == kMutBeerNoDistance) return transmittance;
  Vec3f result;
  float dRef = 1.0f;
  float invdRef = 1.0f / dRef;
  for (int i = 0; i < 3; ++i)
    result[i] = transmittance[i] <= 0.0f ? 0.0f : hm::powf_pos(transmittance[i], invdRef * d);
  return result;
}

// MIS.h:10-21
float msiWeightPowerHeuristic2(int index, const float* pdfs, int n) {
  float s = 0.0f;
  for (int i = 0; i < n; ++i) s += powci<2>(pdfs[i]);
  return powci<2>(pdfs[index]) / s;
}

// MIS.cpp:6-30
MISSampleResult brdfStrategySample(RenderContext ctx, const LightRayInfo& lightRay, const RayHit& hit) {
  MISSampleResult r;
  r.lightRay = lightRay;
  SampleResult s = ctx.scene.sampleBRDF(ctx.rng, r.lightRay, hit, false, kAll);
  if (s.pdfOmega == 0.0f) {
    r.throughputDiffuse = Vec3f::Zero();
    r.throughputConcentrated = Vec3f::Zero();
    r.pdfOmega = 0.0f;
    return r;
  }
  r.ray = ctx.scene.constructRay(hit.interaction, s.direction, s.outside);
  r.throughputDiffuse = s.throughputDiffuse;
  r.throughputConcentrated = s.throughputConcentrated;
  r.pdfOmega = s.pdfOmega;
  ctx.scene.rayHit(r.ray, r.rayhit);
  return r;
}
// MIS.cpp:32-54
MISSampleResult lightStrategySample(RenderContext ctx, const LightRayInfo& lightRay, const Ray& ray,
                                           const RayHit& hit) {
  MISSampleResult r;
  r.lightRay = lightRay;
  bool lightVisible = false;
  r.ray = ctx.scene.sampleLightRayFromStartPoint(ctx.rng, hit.interaction, r.pdfOmega, r.rayhit, lightVisible);
  if (!lightVisible || r.pdfOmega == 0.0f) {
    r.throughputDiffuse = Vec3f::Zero();
    r.throughputConcentrated = Vec3f::Zero();
    r.pdfOmega = 0.0f;
    return r;
  }
  MaterialEvalResult ev =
      ctx.scene.evaluteBRDF(hit, -ray.direction(), r.ray.direction(), r.lightRay.getOutsideIOR(hit), false, kAll);
  float m = std::fabs(dot(r.ray.direction(), hit.interaction.normalGeometric)) / r.pdfOmega;
  r.throughputDiffuse = ev.fDiffuse * m;
  r.throughputConcentrated = ev.fConcentrated * m;
  return r;
}

// PathTracing.cpp:21-95
Vec3f estimateRadiance(RenderContext ctx, LightRayInfo& lightRay, const Ray& ray, Vec3f& normal, Vec3f& albedo,
                       int depth, RayHit* rayhitPtr) {
  if (depth >= ctx.settings.maxDepth) return Vec3f::Zero();
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
      return cwiseProd(throughput, ctx.scene.getEmissionRadiance(-ray.direction(), rayhit));
    }
    // SamplingStrategies (MIS.cpp:60-72): {light = 0, brdf = 1} or {brdf = 0}
    bool twoStrategies = ctx.scene.hasLight() && ctx.settings.useMIS;
    int numStrategies = twoStrategies ? 2 : 1;
    int lightIndex = twoStrategies ? 0 : -1;
    float pdfs[2] = {0.0f, 0.0f};
    float rouletteQ = 0.0f;
    if (depth >= ctx.settings.rouletteStartDepth) rouletteQ = ctx.settings.rouletteQ;
    if (uniform01f(ctx.rng) < rouletteQ) return Vec3f::Zero();
    throughput /= (1.0f - rouletteQ);
    Vec3f estimator = Vec3f::Zero();
    for (int strategy = 0; strategy < numStrategies; ++strategy) {
      bool isLight = strategy == lightIndex;
      MISSampleResult s =
          isLight ? lightStrategySample(ctx, lightRay, ray, rayhit) : brdfStrategySample(ctx, lightRay, rayhit);
      pdfs[strategy] = s.pdfOmega;
      Vec3f throughput2 = cwiseProd(throughput, s.throughput());
      if (throughput2 == Vec3f::Zero() || isSelfIntersection(rayhit, s.rayhit)) continue;
      for (int j = 0; j < numStrategies; ++j) {
        if (j == strategy) continue;
        if (j == lightIndex) pdfs[j] = ctx.scene.evalLightRayFromStartPointDensity(rayhit.interaction, s.ray, s.rayhit);
        else pdfs[j] = ctx.scene.evaluateSamplePDF(rayhit, s.lightRay, s.ray.direction());
      }
      float msiWeight = msiWeightPowerHeuristic2(strategy, pdfs, numStrategies);
      throughput2 *= msiWeight;
      s.lightRay.updateMedia(ctx.scene, rayhit, s.ray.direction());
      Vec3f normal2, albedo2;
      Vec3f Li = estimateRadiance(ctx, s.lightRay, s.ray, normal2, albedo2, depth + 1, &s.rayhit);
      estimator += cwiseProd(Li, throughput2);
    }
    return estimator;
  }
  Vec3f transmittance = beersLaw(lightRay.getTransmittance(), 1e6f);
  Vec3f Le = ctx.scene.evalEnvironment(ray);
  Vec3f L = cwiseProd(transmittance, Le);
  normal = Vec3f::Zero();
  albedo = L;
  return L;
}

// PathTracing.cpp:98-160, one pass per sample index, pixels split over threads.
void renderPT(const PTJob& job, uint32_t firstSample, uint32_t nSamples, int threads, float* color, float* normal,
              float* albedo, uint64_t* nanCount, Bvh2::Stats* stats) {
  const RenderSettings& st = job.settings;
  uint32_t W = st.width, H = st.height;
  if (threads < 1) threads = 1;
  std::atomic<uint64_t> nans{0};
  for (uint32_t k = firstSample; k < firstSample + nSamples; ++k) {
    Vec2f offset = haltonOffset(k);
    auto worker = [&](int tid) {
      for (uint32_t pix = (uint32_t)tid; pix < W * H; pix += (uint32_t)threads) {
        uint32_t i = pix % W, j = pix / W;
        RNG rng = hrng::make_stream(job.seed, j * W + i, k, hrng::kDomainPath);
        RenderContext ctx{job.scene, rng, st};
        Vec2f p0((float)i, (float)j);
        Ray ray = job.scene.computeRay(rng, p0 + offset, (float)W, (float)H);
        Vec3f n, a;
        LightRayInfo info;
        Vec3f L = estimateRadiance(ctx, info, ray, n, a, 0);
        if (L != L) {
          nans++;
          continue;
        }
        size_t o = ((size_t)j * W + i) * 3;
        for (int c = 0; c < 3; ++c) {
          color[o + c] += L[c];
          normal[o + c] += n[c];
          albedo[o + c] += a[c];
        }
      }
    };
    if (stats) {
      // Counting runs single-threaded so the counters need no atomics.
      job.scene.stats = stats;
      worker(0);
      for (int t = 1; t < threads; ++t) worker(t);
      job.scene.stats = nullptr;
    } else if (threads == 1) {
      worker(0);
    } else {
      std::vector<std::thread> pool;
      for (int t = 0; t < threads; ++t) pool.emplace_back(worker, t);
      for (auto& th : pool) th.join();
    }
  }
  if (nanCount) *nanCount += nans.load();
}

}  // namespace oracle

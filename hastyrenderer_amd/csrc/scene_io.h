// Host-side scene loading: renderJob.json -> scene JSON -> .obj/.mtl (+
// environment map and image textures, read by image_io.h), with the reference's key names, defaults and error
// behaviour.  Pure C++17 with no third-party dependency, so the same loader
// runs on every box.  Both the HIP library and the CPU oracle consume the
// resulting `Job`.
//
//   RenderSettings keys/defaults ..... core/src/RenderJob.cpp:10-20, RenderJob.h:21-29
//   RenderJob::loadJSON .............. core/src/RenderJob.cpp:21-28
//   Camera from_json / setLookAt ..... core/src/Camera.cpp:9-39
//   Scene::Scene ..................... core/src/Scene.cpp:12-195
//   tinyobj collectTriangle* ......... core/src/tinyobj.cpp:7-88
//   EnvironmentTexture::load ......... core/include/Hasty/Environment.h:105-118
//   Texture3f/1f::loadTexture ........ core/src/Texture.cpp:21-49, Image.cpp:74-138
//   ImageDistribution::precompute .... core/include/Hasty/Environment.h:52-58
#pragma once
#include <stdint.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hasty {

// ---------------------------------------------------------------- JSON ----
struct JsonValue {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  bool isInteger = false;
  std::string str;
  std::vector<JsonValue> arr;
  std::vector<std::pair<std::string, JsonValue>> obj;

  bool contains(const std::string& key) const;
  const JsonValue& at(const std::string& key) const;  // throws if missing
  const JsonValue& at(size_t i) const;                // throws if out of range
  float asFloat() const;
  double asDouble() const;
  int asInt() const;
  uint32_t asUInt() const;
  bool asBool() const;
  const std::string& asString() const;
};
JsonValue parseJson(const std::string& text);
JsonValue readJsonFile(const std::string& path);

// ------------------------------------------------------------ settings ----
struct RenderSettings {
  int rouletteStartDepth = 1;
  float rouletteQ = 0.3f;
  uint32_t width = 720;
  uint32_t height = 720;
  uint32_t numThreads = 4;
  int maxDepth = 30;
  int numSamples = 32;
  bool useMIS = true;
  float exposure = 0.0f;  // declared by the reference but never parsed
};
void settingsFromJson(const JsonValue& j, RenderSettings& s);

// -------------------------------------------------------------- camera ----
struct Camera {
  float position[3] = {0, 0, 0};
  float forward[3] = {0, 0, -1};
  float right[3] = {1, 0, 0};
  float up[3] = {0, 1, 0};
  float fovDegree = 45.0f;  // fixed in the reference; optional "fov" key here
  float aperture = 0.0f;
  float focalDistance = 1.0f;
  float exposure = 0.0f;
  int numBlades = 8;
  float bladeRotation = 0.0f;
  void setLookAt(const float pos[3], const float lookAt[3]);
};

// ------------------------------------------------------------ material ----
// PrincipledBRDF parameters after the MTL mapping of Scene.cpp:65-118.
struct Material {
  std::string name;
  float albedo[3] = {0, 0, 0};  // Kd
  float roughness = 0.0f;       // 1 - sqrt(Ns / 1000)
  float metallic = 0.0f;        // Ka[0]
  float specular = 0.0f;        // Ks[0]
  float ior = 1.0f;             // Ni
  float anisotropy = 0.0f;      // aniso
  float transmission = 0.0f;    // 1 - d
  bool varyIOR = false;         // world.vary_glass_ior && transmission > 0
  float emission[3] = {0, 0, 0};  // Ke
  // Image textures (indices into Scene::textures, -1 = the constant above):
  // map_Kd (sRGB, Texture3f), map_Ns (roughness, Texture1f), map_Pm
  // (metallic, Texture1f) and bump (normal map, Texture3f), Scene.cpp:73-114.
  int texAlbedo = -1, texRoughness = -1, texMetallic = -1, texNormal = -1;
  bool isLight() const { return emission[0] > 0.0f || emission[1] > 0.0f || emission[2] > 0.0f; }
  bool textured() const { return texAlbedo >= 0 || texRoughness >= 0 || texMetallic >= 0 || texNormal >= 0; }
};

// An image texture after TextureTraits::loadTexture (Texture.cpp:21-49):
// linear float texels, row 0 = top, 1 channel (Texture1f) or 3 (Texture3f;
// the alpha of a 4-channel file is dropped as readImage3f drops it).
struct Texture {
  std::string path;
  uint32_t width = 0, height = 0, channels = 0;
  std::vector<float> px;
};
// Reads an image file as a texture of `channels` (1 or 3) channels, with the
// reference's channel-count errors (Image.cpp:74-138); srgb applies
// pow(clamp(p, 0, FLT_MAX), 2.2) per channel (Texture.cpp:28, 43).
Texture loadTexture(const std::string& path, uint32_t channels, bool srgb);

// --------------------------------------------------------------- image ----
struct ImageRGB {
  uint32_t width = 0, height = 0;
  std::vector<float> rgb;  // row-major, row 0 = top (OIIO in-memory order)
};
ImageRGB readPFM(const std::string& path);
// Any image file readImage understands (PFM, PPM/PGM, PNG, EXR) with 3 or 4
// channels, as readImage3f requires (Image.cpp:104-138); alpha is dropped.
ImageRGB readImageRGB(const std::string& path);
// Writes a colour PFM following the format's bottom-to-top row order.
void writePFM(const std::string& path, uint32_t width, uint32_t height, const float* rgbTopRowFirst);

// --------------------------------------------------------------- scene ----
struct Scene {
  Camera camera;
  bool hasEnvTexture = false;
  float backgroundColor[3] = {0, 0, 0};
  bool varyGlassIOR = false;

  // Environment map (only when hasEnvTexture).
  ImageRGB env;
  std::vector<float> envPdf;   // norm(pixel) * sin(theta_row), W*H
  float envValueSum = 0.0f;    // float accumulation in memory order
  std::vector<double> envCdf;  // discrete CDF over envPdf, last entry 1

  // tinyobj-equivalent attribute arrays and shapes.
  std::vector<float> vertices, normals, texcoords;
  std::vector<Material> materials;
  std::vector<Texture> textures;
  std::vector<std::string> shapeNames;
  std::vector<uint32_t> shapeTriOffset;  // global triangle index = offset[geomID] + primID

  // Flattened triangles, indexed by global triangle index.
  uint32_t numTris = 0;
  std::vector<float> triPos;       // 9 per triangle: p0, p1, p2 (collectTriangle)
  std::vector<float> triNrm;       // 9 per triangle (collectTriangleNormals; valid if smooth)
  std::vector<float> triUV;        // 6 per triangle (valid if hasUV)
  std::vector<uint8_t> triSmooth;  // smoothing group != 0
  std::vector<uint8_t> triHasUV;
  std::vector<int32_t> triMaterial;
  std::vector<uint32_t> triGeomID, triPrimID;

  // Surface lights (Scene.cpp:159-181, 193-194).
  std::vector<uint32_t> lightTris;  // global triangle indices, file order
  std::vector<float> lightAreas;
  float totalLightArea = 0.0f;
  std::vector<double> lightCdf;

  bool hasSurfaceLight() const { return !lightTris.empty(); }
  bool backgroundIsZero() const {
    return !hasEnvTexture && backgroundColor[0] == 0.0f && backgroundColor[1] == 0.0f &&
           backgroundColor[2] == 0.0f;
  }
  bool hasLight() const { return hasSurfaceLight() || !backgroundIsZero(); }
};

struct Job {
  RenderSettings settings;
  Scene scene;
  std::string jobPath;
};

// Faces of one shape as tinyobj lists them (indices into the attribute
// arrays, -1 = absent), one material id and smoothing group per triangle.
struct FaceIndex {
  int v, t, n;
};
struct ShapeFaces {
  std::string name;
  std::vector<FaceIndex> indices;  // 3 per triangle
  std::vector<int> materialIds;
  std::vector<uint32_t> smoothing;
};

// The environment map and its sampling tables (Environment.h:105-118,
// 52-58); throws if the image has no energy.
void setEnvironment(Scene& sc, ImageRGB env);

// Flattens the shapes' triangles into the Scene arrays (collectTriangle*,
// tinyobj.cpp:7-88) and collects the surface lights (Scene.cpp:159-181,
// 193-194).  sc.vertices / normals / texcoords / materials must be set;
// emitterTinted[m] marks an emissive material with a non-black Kd or Ks,
// which the reference refuses when a face uses it.
void buildTriangles(Scene& sc, const std::vector<ShapeFaces>& shapes, const std::vector<uint8_t>& emitterTinted);

// Throws std::runtime_error with the reference's messages on malformed input.
void loadScene(const std::string& sceneJsonPath, Scene& out);
void loadJob(const std::string& renderJobPath, Job& out);

// Discrete CDF (replacement for std::discrete_distribution's table): p_i =
// w_i / sum(w) in double, partial sums, last entry forced to 1.
std::vector<double> buildDiscreteCdf(const float* w, size_t n);

}  // namespace hasty

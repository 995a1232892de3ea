// Scene loading for the renderer; see scene_io.h for the reference mapping.
#include "scene_io.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <fstream>
#include <sstream>

#include "hmath.h"
#include "image_io.h"

namespace hasty {

// =================================================================== JSON ===
namespace {

struct JsonParser {
  const std::string& s;
  size_t p = 0;
  explicit JsonParser(const std::string& text) : s(text) {}

  [[noreturn]] void fail(const char* what) {
    throw std::runtime_error(std::string("json parse error: ") + what + " at offset " + std::to_string(p));
  }
  void ws() {
    while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p;
  }
  JsonValue value() {
    ws();
    if (p >= s.size()) fail("unexpected end");
    char c = s[p];
    if (c == '{') return object();
    if (c == '[') return array();
    if (c == '"') {
      JsonValue v;
      v.kind = JsonValue::String;
      v.str = string();
      return v;
    }
    if (c == 't' || c == 'f' || c == 'n') return literal();
    return number();
  }
  JsonValue literal() {
    JsonValue v;
    if (s.compare(p, 4, "true") == 0) {
      v.kind = JsonValue::Bool;
      v.b = true;
      p += 4;
    } else if (s.compare(p, 5, "false") == 0) {
      v.kind = JsonValue::Bool;
      v.b = false;
      p += 5;
    } else if (s.compare(p, 4, "null") == 0) {
      v.kind = JsonValue::Null;
      p += 4;
    } else {
      fail("bad literal");
    }
    return v;
  }
  JsonValue number() {
    size_t start = p;
    bool integer = true;
    if (p < s.size() && (s[p] == '-' || s[p] == '+')) ++p;
    while (p < s.size()) {
      char c = s[p];
      if (c >= '0' && c <= '9') {
        ++p;
      } else if (c == '.' || c == 'e' || c == 'E' || c == '-' || c == '+') {
        integer = false;
        ++p;
      } else {
        break;
      }
    }
    if (p == start) fail("bad value");
    std::string tok = s.substr(start, p - start);
    char* end = nullptr;
    double d = strtod(tok.c_str(), &end);
    if (end == tok.c_str() || *end != '\0') fail("bad number");
    JsonValue v;
    v.kind = JsonValue::Number;
    v.num = d;
    v.isInteger = integer;
    return v;
  }
  std::string string() {
    if (s[p] != '"') fail("expected string");
    ++p;
    std::string out;
    while (p < s.size() && s[p] != '"') {
      char c = s[p++];
      if (c == '\\') {
        if (p >= s.size()) fail("bad escape");
        char e = s[p++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (p + 4 > s.size()) fail("bad unicode escape");
            unsigned code = (unsigned)strtoul(s.substr(p, 4).c_str(), nullptr, 16);
            p += 4;
            if (code < 0x80) {
              out += (char)code;
            } else if (code < 0x800) {
              out += (char)(0xC0 | (code >> 6));
              out += (char)(0x80 | (code & 0x3F));
            } else {
              out += (char)(0xE0 | (code >> 12));
              out += (char)(0x80 | ((code >> 6) & 0x3F));
              out += (char)(0x80 | (code & 0x3F));
            }
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (p >= s.size()) fail("unterminated string");
    ++p;
    return out;
  }
  JsonValue array() {
    JsonValue v;
    v.kind = JsonValue::Array;
    ++p;
    ws();
    if (p < s.size() && s[p] == ']') {
      ++p;
      return v;
    }
    while (true) {
      v.arr.push_back(value());
      ws();
      if (p >= s.size()) fail("unterminated array");
      if (s[p] == ',') {
        ++p;
        continue;
      }
      if (s[p] == ']') {
        ++p;
        break;
      }
      fail("expected , or ]");
    }
    return v;
  }
  JsonValue object() {
    JsonValue v;
    v.kind = JsonValue::Object;
    ++p;
    ws();
    if (p < s.size() && s[p] == '}') {
      ++p;
      return v;
    }
    while (true) {
      ws();
      std::string key = string();
      ws();
      if (p >= s.size() || s[p] != ':') fail("expected :");
      ++p;
      JsonValue val = value();
      v.obj.emplace_back(std::move(key), std::move(val));
      ws();
      if (p >= s.size()) fail("unterminated object");
      if (s[p] == ',') {
        ++p;
        continue;
      }
      if (s[p] == '}') {
        ++p;
        break;
      }
      fail("expected , or }");
    }
    return v;
  }
};

std::string readTextFile(const std::string& path) {
  std::ifstream ifs(path, std::ios::binary);
  if (!ifs.is_open()) {
    // Json.cpp:12-15
    throw std::runtime_error(std::string("could not open ") + path + "  for reading ");
  }
  std::stringstream ss;
  ss << ifs.rdbuf();
  return ss.str();
}

std::string parentDir(const std::string& path) {
  size_t slash = path.find_last_of('/');
  if (slash == std::string::npos) return ".";
  if (slash == 0) return "/";
  return path.substr(0, slash);
}
std::string joinPath(const std::string& dir, const std::string& rel) {
  if (!rel.empty() && rel[0] == '/') return rel;
  if (dir.empty() || dir == ".") return rel;
  return dir + "/" + rel;
}

}  // namespace

bool JsonValue::contains(const std::string& key) const {
  if (kind != Object) return false;
  for (auto& kv : obj)
    if (kv.first == key) return true;
  return false;
}
const JsonValue& JsonValue::at(const std::string& key) const {
  if (kind == Object) {
    for (auto& kv : obj)
      if (kv.first == key) return kv.second;
  }
  throw std::runtime_error("json key '" + key + "' not found");
}
const JsonValue& JsonValue::at(size_t i) const {
  if (kind != Array || i >= arr.size()) throw std::runtime_error("json array index out of range");
  return arr[i];
}
double JsonValue::asDouble() const {
  if (kind != Number) throw std::runtime_error("json value is not a number");
  return num;
}
float JsonValue::asFloat() const { return (float)asDouble(); }
int JsonValue::asInt() const { return (int)asDouble(); }
uint32_t JsonValue::asUInt() const { return (uint32_t)asDouble(); }
bool JsonValue::asBool() const {
  if (kind != Bool) throw std::runtime_error("json value is not a boolean");
  return b;
}
const std::string& JsonValue::asString() const {
  if (kind != String) throw std::runtime_error("json value is not a string");
  return str;
}

JsonValue parseJson(const std::string& text) {
  JsonParser parser(text);
  JsonValue v = parser.value();
  parser.ws();
  if (parser.p != text.size()) parser.fail("trailing characters");
  return v;
}
JsonValue readJsonFile(const std::string& path) { return parseJson(readTextFile(path)); }

// RenderJob.cpp:10-20
void settingsFromJson(const JsonValue& j, RenderSettings& s) {
  s.rouletteStartDepth = j.contains("roulette_start_depth") ? j.at("roulette_start_depth").asInt() : 1;
  s.rouletteQ = j.contains("roulette_q") ? j.at("roulette_q").asFloat() : 0.3f;
  s.width = j.contains("width") ? j.at("width").asUInt() : 720u;
  s.height = j.contains("height") ? j.at("height").asUInt() : 720u;
  s.numThreads = j.contains("num_threads") ? j.at("num_threads").asUInt() : 4u;
  s.maxDepth = j.contains("max_depth") ? j.at("max_depth").asInt() : 30;
  s.numSamples = j.contains("num_samples") ? j.at("num_samples").asInt() : 32;
  s.useMIS = j.contains("use_mis") ? j.at("use_mis").asBool() : true;
}

// ================================================================= camera ===
static void vec3FromJson(const JsonValue& j, float out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = j.at((size_t)i).asFloat();  // VMath.cpp:10-16
}

// Camera.cpp:31-39
void Camera::setLookAt(const float pos[3], const float lookAt[3]) {
  for (int i = 0; i < 3; ++i) position[i] = pos[i];
  float f[3] = {lookAt[0] - pos[0], lookAt[1] - pos[1], lookAt[2] - pos[2]};
  float n = hm::fsqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (int i = 0; i < 3; ++i) forward[i] = f[i] / n;
  float u[3] = {0.0f, 1.0f, 0.0f};
  float d = u[0] * forward[0] + u[1] * forward[1] + u[2] * forward[2];
  for (int i = 0; i < 3; ++i) u[i] = u[i] - forward[i] * d;
  float un = hm::fsqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int i = 0; i < 3; ++i) up[i] = u[i] / un;
  right[0] = forward[1] * up[2] - forward[2] * up[1];
  right[1] = forward[2] * up[0] - forward[0] * up[2];
  right[2] = forward[0] * up[1] - forward[1] * up[0];
}

// Camera.cpp:9-25 (+ optional "fov", degrees, default 45 as Camera.h:41)
static void cameraFromJson(const JsonValue& j, Camera& c) {
  float pos[3];
  vec3FromJson(j.at("position"), pos);
  float look[3];
  if (j.contains("look_at")) {
    vec3FromJson(j.at("look_at"), look);
  } else {
    look[0] = pos[0] + 0.0f;
    look[1] = pos[1] + 0.0f;
    look[2] = pos[2] + -1.0f;
  }
  c.setLookAt(pos, look);
  c.aperture = j.contains("aperture") ? j.at("aperture").asFloat() : 0.0f;
  c.focalDistance = j.contains("focal_distance") ? j.at("focal_distance").asFloat() : 1.0f;
  c.exposure = j.contains("exposure") ? j.at("exposure").asFloat() : 0.0f;
  c.numBlades = j.contains("num_blades") ? j.at("num_blades").asInt() : 8;
  c.bladeRotation = j.contains("blade_rotation") ? j.at("blade_rotation").asFloat() : 0.0f;
  c.fovDegree = j.contains("fov") ? j.at("fov").asFloat() : 45.0f;
  if (c.aperture > 0.0f && c.numBlades < 3) throw std::runtime_error("camera num_blades must be >= 3");
}

// ================================================================== image ===
ImageRGB readPFM(const std::string& path) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("could not open " + path + " for reading");
  char magic[3] = {0, 0, 0};
  int width = 0, height = 0;
  double scale = 0.0;
  if (fscanf(f, "%2s %d %d %lf", magic, &width, &height, &scale) != 4 || width <= 0 || height <= 0) {
    fclose(f);
    throw std::runtime_error("malformed PFM header in " + path);
  }
  fgetc(f);  // the single whitespace byte after the scale
  int channels = 0;
  if (strcmp(magic, "PF") == 0) channels = 3;
  else if (strcmp(magic, "Pf") == 0) channels = 1;
  else {
    fclose(f);
    throw std::runtime_error("not a PFM file: " + path);
  }
  size_t count = (size_t)width * height * channels;
  std::vector<float> raw(count);
  if (fread(raw.data(), sizeof(float), count, f) != count) {
    fclose(f);
    throw std::runtime_error("truncated PFM data in " + path);
  }
  fclose(f);
  if (scale > 0.0) {  // big-endian payload
    for (float& v : raw) {
      uint32_t u;
      memcpy(&u, &v, 4);
      u = __builtin_bswap32(u);
      memcpy(&v, &u, 4);
    }
  }
  ImageRGB img;
  img.width = (uint32_t)width;
  img.height = (uint32_t)height;
  img.rgb.resize((size_t)width * height * 3);
  for (int y = 0; y < height; ++y) {
    int srcRow = height - 1 - y;  // PFM stores the bottom row first
    for (int x = 0; x < width; ++x) {
      for (int c = 0; c < 3; ++c) {
        float v = raw[((size_t)srcRow * width + x) * channels + (channels == 3 ? c : 0)];
        img.rgb[((size_t)y * width + x) * 3 + c] = v;
      }
    }
  }
  return img;
}

ImageRGB readImageRGB(const std::string& path) {
  FloatImage im = readImage(path);
  if (im.channels != 3 && im.channels != 4)
    throw std::runtime_error("could not read " + path + ", wrong number of channels, expected 3 or 4, found " +
                             std::to_string(im.channels));
  ImageRGB out;
  out.width = im.width;
  out.height = im.height;
  out.rgb.resize((size_t)im.width * im.height * 3);
  for (size_t i = 0; i < (size_t)im.width * im.height; ++i)
    for (int c = 0; c < 3; ++c) out.rgb[3 * i + c] = im.px[i * im.channels + c];
  return out;
}

Texture loadTexture(const std::string& path, uint32_t channels, bool srgb) {
  Texture t;
  t.path = path;
  t.channels = channels;
  if (channels == 3) {
    ImageRGB im = readImageRGB(path);
    t.width = im.width;
    t.height = im.height;
    t.px = std::move(im.rgb);
  } else if (channels == 1) {
    FloatImage im = readImage(path);
    if (im.channels != 1)
      throw std::runtime_error("could not read " + path + ", wrong number of channels, expected 1, found " +
                               std::to_string(im.channels));
    t.width = im.width;
    t.height = im.height;
    t.px = std::move(im.px);
  } else {
    throw std::runtime_error("textures have 1 or 3 channels");
  }
  if (srgb)  // clampAndPowInplace(image, 0, FLT_MAX, 2.2) (Image.h:320-326)
    for (float& v : t.px) v = std::pow(std::min(std::max(v, 0.0f), std::numeric_limits<float>::max()), 2.2f);
  return t;
}

void writePFM(const std::string& path, uint32_t width, uint32_t height, const float* rgb) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("could not open file " + path + "for writing");
  fprintf(f, "PF\n%u %u\n-1.0\n", width, height);
  for (uint32_t y = 0; y < height; ++y) {
    const float* row = rgb + (size_t)(height - 1 - y) * width * 3;
    fwrite(row, sizeof(float), (size_t)width * 3, f);
  }
  fclose(f);
}

// ============================================================ OBJ / MTL ===
namespace {

using ObjIndex = FaceIndex;
using ObjShape = ShapeFaces;

struct RawMaterial {
  std::string name;
  float ambient[3] = {0, 0, 0}, diffuse[3] = {0, 0, 0}, specular[3] = {0, 0, 0};
  float emission[3] = {0, 0, 0};
  float shininess = 1.0f, ior = 1.0f, dissolve = 1.0f, anisotropy = 0.0f;
  std::string diffuseTex, specularHighlightTex, metallicTex, bumpTex;
};

std::vector<std::string> splitWs(const std::string& line) {
  std::vector<std::string> out;
  size_t i = 0, n = line.size();
  while (i < n) {
    while (i < n && (line[i] == ' ' || line[i] == '\t' || line[i] == '\r')) ++i;
    if (i >= n) break;
    size_t j = i;
    while (j < n && line[j] != ' ' && line[j] != '\t' && line[j] != '\r') ++j;
    out.push_back(line.substr(i, j - i));
    i = j;
  }
  return out;
}
float parseFloat(const std::string& t) { return (float)strtod(t.c_str(), nullptr); }
void parseReal3(const std::vector<std::string>& tok, float out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = (size_t)(i + 1) < tok.size() ? parseFloat(tok[i + 1]) : 0.0f;
}
std::string restOfLine(const std::string& line, const std::string& key) {
  size_t pos = line.find(key);
  pos += key.size();
  while (pos < line.size() && (line[pos] == ' ' || line[pos] == '\t')) ++pos;
  size_t end = line.size();
  while (end > pos && (line[end - 1] == ' ' || line[end - 1] == '\r' || line[end - 1] == '\t')) --end;
  return line.substr(pos, end - pos);
}

// The file name of a texture statement: tinyobj's ParseTextureNameAndOption
// skips the option flags (-bm 1.0, -o u v w, -clamp on, ...) and takes the
// rest of the line as the name.
std::string textureName(const std::string& line, const std::string& key) {
  static const std::map<std::string, int> optionArgs = {
      {"-blendu", 1}, {"-blendv", 1}, {"-clamp", 1}, {"-boost", 1}, {"-bm", 1}, {"-mm", 2},
      {"-o", 3},      {"-s", 3},      {"-t", 3},     {"-texres", 1}, {"-imfchan", 1}, {"-type", 1},
      {"-colorspace", 1}};
  std::string rest = restOfLine(line, key);
  for (;;) {
    auto tok = splitWs(rest);
    if (tok.empty()) return std::string();
    auto it = optionArgs.find(tok[0]);
    if (it == optionArgs.end()) return rest;
    size_t skip = 1 + (size_t)it->second;
    size_t pos = 0;
    for (size_t k = 0; k < skip && pos < rest.size(); ++k) {
      while (pos < rest.size() && (rest[pos] == ' ' || rest[pos] == '\t')) ++pos;
      while (pos < rest.size() && rest[pos] != ' ' && rest[pos] != '\t') ++pos;
    }
    while (pos < rest.size() && (rest[pos] == ' ' || rest[pos] == '\t')) ++pos;
    rest = rest.substr(pos);
  }
}

void loadMtl(const std::string& path, std::vector<RawMaterial>& mats, std::map<std::string, int>& byName) {
  std::ifstream ifs(path);
  if (!ifs.is_open()) throw std::runtime_error("TinyObjReader error could not open material file " + path);
  std::string line;
  RawMaterial* cur = nullptr;
  while (std::getline(ifs, line)) {
    auto tok = splitWs(line);
    if (tok.empty() || tok[0][0] == '#') continue;
    const std::string& k = tok[0];
    if (k == "newmtl") {
      RawMaterial m;
      m.name = tok.size() > 1 ? restOfLine(line, "newmtl") : std::string();
      byName[m.name] = (int)mats.size();
      mats.push_back(m);
      cur = &mats.back();
      continue;
    }
    if (!cur) continue;
    if (k == "Ka") parseReal3(tok, cur->ambient);
    else if (k == "Kd") parseReal3(tok, cur->diffuse);
    else if (k == "Ks") parseReal3(tok, cur->specular);
    else if (k == "Ke") parseReal3(tok, cur->emission);
    else if (k == "Ns" && tok.size() > 1) cur->shininess = parseFloat(tok[1]);
    else if (k == "Ni" && tok.size() > 1) cur->ior = parseFloat(tok[1]);
    else if (k == "d" && tok.size() > 1) cur->dissolve = parseFloat(tok[1]);
    else if (k == "Tr" && tok.size() > 1) cur->dissolve = 1.0f - parseFloat(tok[1]);
    else if (k == "aniso" && tok.size() > 1) cur->anisotropy = parseFloat(tok[1]);
    else if (k == "map_Kd") cur->diffuseTex = textureName(line, "map_Kd");
    else if (k == "map_Ns") cur->specularHighlightTex = textureName(line, "map_Ns");
    else if (k == "map_Pm") cur->metallicTex = textureName(line, "map_Pm");
    else if (k == "bump" || k == "map_bump" || k == "map_Bump") cur->bumpTex = textureName(line, k);
  }
}

int resolveIndex(int idx, size_t count) {
  if (idx > 0) return idx - 1;
  if (idx < 0) return (int)count + idx;
  return -1;
}

void parseObj(const std::string& path, const std::string& mtlSearchDir, std::vector<float>& vertices,
              std::vector<float>& normals, std::vector<float>& texcoords, std::vector<ObjShape>& shapes,
              std::vector<RawMaterial>& materials) {
  std::ifstream ifs(path);
  if (!ifs.is_open()) throw std::runtime_error("TinyObjReader error Cannot open file [" + path + "]");
  std::map<std::string, int> matByName;
  ObjShape cur;
  int currentMaterial = -1;
  uint32_t currentSmoothing = 0;
  auto flush = [&]() {
    if (!cur.materialIds.empty()) shapes.push_back(cur);
  };
  std::string line;
  while (std::getline(ifs, line)) {
    auto tok = splitWs(line);
    if (tok.empty() || tok[0][0] == '#') continue;
    const std::string& k = tok[0];
    if (k == "v") {
      for (int i = 1; i <= 3; ++i) vertices.push_back(i < (int)tok.size() ? parseFloat(tok[i]) : 0.0f);
    } else if (k == "vn") {
      for (int i = 1; i <= 3; ++i) normals.push_back(i < (int)tok.size() ? parseFloat(tok[i]) : 0.0f);
    } else if (k == "vt") {
      for (int i = 1; i <= 2; ++i) texcoords.push_back(i < (int)tok.size() ? parseFloat(tok[i]) : 0.0f);
    } else if (k == "f") {
      std::vector<ObjIndex> poly;
      for (size_t i = 1; i < tok.size(); ++i) {
        ObjIndex ix{-1, -1, -1};
        const std::string& t = tok[i];
        size_t s1 = t.find('/');
        if (s1 == std::string::npos) {
          ix.v = resolveIndex(atoi(t.c_str()), vertices.size() / 3);
        } else {
          ix.v = resolveIndex(atoi(t.substr(0, s1).c_str()), vertices.size() / 3);
          size_t s2 = t.find('/', s1 + 1);
          std::string ts = t.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
          // "f 1/5" or "f 1//9" naming more vt / vn lines than the file has
          if (!ts.empty()) {
            ix.t = resolveIndex(atoi(ts.c_str()), texcoords.size() / 2);
            if (ix.t < 0 || (size_t)ix.t >= texcoords.size() / 2)
              throw std::runtime_error("TinyObjReader error invalid texcoord index in " + path);
          }
          if (s2 != std::string::npos) {
            std::string ns = t.substr(s2 + 1);
            if (!ns.empty()) {
              ix.n = resolveIndex(atoi(ns.c_str()), normals.size() / 3);
              if (ix.n < 0 || (size_t)ix.n >= normals.size() / 3)
                throw std::runtime_error("TinyObjReader error invalid normal index in " + path);
            }
          }
        }
        if (ix.v < 0 || (size_t)ix.v >= vertices.size() / 3)
          throw std::runtime_error("TinyObjReader error invalid vertex index in " + path);
        poly.push_back(ix);
      }
      if (poly.size() < 3) continue;
      // Fan triangulation of polygons (triangles pass through unchanged).
      for (size_t i = 1; i + 1 < poly.size(); ++i) {
        cur.indices.push_back(poly[0]);
        cur.indices.push_back(poly[i]);
        cur.indices.push_back(poly[i + 1]);
        cur.materialIds.push_back(currentMaterial);
        cur.smoothing.push_back(currentSmoothing);
      }
    } else if (k == "o" || k == "g") {
      flush();
      cur = ObjShape();
      cur.name = tok.size() > 1 ? restOfLine(line, k) : std::string();
    } else if (k == "s") {
      if (tok.size() > 1 && tok[1] != "off") currentSmoothing = (uint32_t)atoi(tok[1].c_str());
      else currentSmoothing = 0;
    } else if (k == "usemtl") {
      std::string name = restOfLine(line, "usemtl");
      auto it = matByName.find(name);
      currentMaterial = it == matByName.end() ? -1 : it->second;
    } else if (k == "mtllib") {
      for (size_t i = 1; i < tok.size(); ++i) loadMtl(joinPath(mtlSearchDir, tok[i]), materials, matByName);
    }
  }
  flush();
}

}  // namespace

std::vector<double> buildDiscreteCdf(const float* w, size_t n) {
  std::vector<double> cdf(n);
  if (n == 0) return cdf;
  double sum = 0.0;
  for (size_t i = 0; i < n; ++i) sum += (double)w[i];
  double acc = 0.0;
  for (size_t i = 0; i < n; ++i) {
    acc += (double)w[i] / sum;
    cdf[i] = acc;
  }
  cdf[n - 1] = 1.0;
  return cdf;
}

// Environment.h:105-118: weights norm(pixel) * sin(theta_row)
void setEnvironment(Scene& sc, ImageRGB env) {
  sc.hasEnvTexture = true;
  sc.env = std::move(env);
  uint32_t W = sc.env.width, H = sc.env.height;
  if (W == 0 || H == 0 || sc.env.rgb.size() != (size_t)W * H * 3)
    throw std::runtime_error("environment texture has no pixels");
  float heightF = (float)H;
  sc.envPdf.assign((size_t)W * H, 0.0f);
  for (uint32_t y = 0; y < H; ++y) {
    float theta = (((float)(int)y + 0.5f) / heightF) * (float)hm::kPi;
    float st = hm::sinf(theta);
    for (uint32_t x = 0; x < W; ++x) {
      const float* px = &sc.env.rgb[((size_t)y * W + x) * 3];
      float n = hm::fsqrt(px[0] * px[0] + px[1] * px[1] + px[2] * px[2]);
      sc.envPdf[(size_t)y * W + x] = n * st;
    }
  }
  float s = 0.0f;
  for (float v : sc.envPdf) s += v;  // std::accumulate(..., 0.f)
  sc.envValueSum = s;
  if (!(s > 0.0f)) throw std::runtime_error("environment texture has no energy");
  sc.envCdf = buildDiscreteCdf(sc.envPdf.data(), sc.envPdf.size());
}

void buildTriangles(Scene& sc, const std::vector<ShapeFaces>& shapes, const std::vector<uint8_t>& emitterTinted) {
  size_t total = 0;
  for (auto& s : shapes) total += s.materialIds.size();
  if (total >= 0x80000000ull) throw std::runtime_error("more than 2^31 - 1 triangles");
  sc.numTris = (uint32_t)total;
  sc.triPos.resize(total * 9);
  sc.triNrm.resize(total * 9);
  sc.triUV.resize(total * 6);
  sc.triSmooth.resize(total);
  sc.triHasUV.resize(total);
  sc.triMaterial.resize(total);
  sc.triGeomID.resize(total);
  sc.triPrimID.resize(total);
  sc.shapeNames.clear();
  sc.shapeTriOffset.clear();
  sc.lightTris.clear();
  sc.lightAreas.clear();
  const size_t nv = sc.vertices.size() / 3, nn = sc.normals.size() / 3, nt = sc.texcoords.size() / 2;
  uint32_t g = 0;
  for (size_t s = 0; s < shapes.size(); ++s) {
    const ShapeFaces& sh = shapes[s];
    sc.shapeNames.push_back(sh.name);
    sc.shapeTriOffset.push_back(g);
    size_t nf = sh.materialIds.size();
    if (sh.indices.size() != 3 * nf || sh.smoothing.size() != nf)
      throw std::runtime_error("shape '" + sh.name + "': index, material and smoothing counts disagree");
    for (size_t f = 0; f < nf; ++f, ++g) {
      int matId = sh.materialIds[f];
      if (matId == -1) throw std::runtime_error(" error face has no material ");
      if (matId < 0 || (size_t)matId >= sc.materials.size()) throw std::runtime_error("material index out of range");
      for (int v = 0; v < 3; ++v) {
        const FaceIndex& ix = sh.indices[3 * f + v];
        if (ix.v < 0 || (size_t)ix.v >= nv) throw std::runtime_error("vertex index out of range");
        if (ix.n >= 0 && (size_t)ix.n >= nn) throw std::runtime_error("normal index out of range");
        if (ix.t >= 0 && (size_t)ix.t >= nt) throw std::runtime_error("texcoord index out of range");
      }
      // collectTriangle (tinyobj.cpp:7-28)
      float p[3][3];
      for (int v = 0; v < 3; ++v) {
        const FaceIndex& ix = sh.indices[3 * f + v];
        for (int c = 0; c < 3; ++c) p[v][c] = sc.vertices[3 * (size_t)ix.v + c];
      }
      for (int v = 0; v < 3; ++v)
        for (int c = 0; c < 3; ++c) sc.triPos[(size_t)g * 9 + v * 3 + c] = p[v][c];
      // collectTriangleNormals (tinyobj.cpp:30-58)
      bool allNormals = true;
      for (int v = 0; v < 3; ++v)
        if (sh.indices[3 * f + v].n < 0) allNormals = false;
      if (allNormals) {
        for (int v = 0; v < 3; ++v) {
          const FaceIndex& ix = sh.indices[3 * f + v];
          for (int c = 0; c < 3; ++c) sc.triNrm[(size_t)g * 9 + v * 3 + c] = sc.normals[3 * (size_t)ix.n + c];
        }
      } else {
        float e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        float e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
        float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        float n = hm::fsqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        for (int v = 0; v < 3; ++v)
          for (int c = 0; c < 3; ++c) sc.triNrm[(size_t)g * 9 + v * 3 + c] = cr[c] / n;
      }
      // collectTriangleUV (tinyobj.cpp:73-88)
      bool hasUV = true;
      for (int v = 0; v < 3; ++v) {
        int t = sh.indices[3 * f + v].t;
        if (t < 0 || 2 * (size_t)t + 1 >= sc.texcoords.size()) hasUV = false;
      }
      sc.triHasUV[g] = hasUV;
      for (int v = 0; v < 3; ++v)
        for (int c = 0; c < 2; ++c)
          sc.triUV[(size_t)g * 6 + v * 2 + c] = hasUV ? sc.texcoords[2 * (size_t)sh.indices[3 * f + v].t + c] : 0.0f;
      sc.triSmooth[g] = sh.smoothing[f] != 0;
      sc.triMaterial[g] = matId;
      sc.triGeomID[g] = (uint32_t)s;
      sc.triPrimID[g] = (uint32_t)f;

      // Scene.cpp:159-181
      if (sc.materials[matId].isLight()) {
        if ((size_t)matId < emitterTinted.size() && emitterTinted[matId])
          throw std::runtime_error(" don't support non black diffuse materials with emission ");
        sc.lightTris.push_back(g);
        float e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        float e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
        float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        float area = 0.5f * hm::fsqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        sc.lightAreas.push_back(area);
      }
    }
  }
  float la = 0.0f;
  for (float a : sc.lightAreas) la += a;  // std::accumulate(..., 0.f)
  sc.totalLightArea = la;
  sc.lightCdf.clear();
  if (!sc.lightAreas.empty()) sc.lightCdf = buildDiscreteCdf(sc.lightAreas.data(), sc.lightAreas.size());
}

// Scene.cpp:12-195
void loadScene(const std::string& sceneJsonPath, Scene& sc) {
  std::string folder = parentDir(sceneJsonPath);
  JsonValue j = readJsonFile(sceneJsonPath);
  cameraFromJson(j.at("camera"), sc.camera);
  const JsonValue& world = j.at("world");
  if (world.contains("background_color")) {
    sc.hasEnvTexture = false;
    vec3FromJson(world.at("background_color"), sc.backgroundColor);
  } else {
    setEnvironment(sc, readImageRGB(joinPath(folder, world.at("environment_texture").asString())));
  }
  sc.varyGlassIOR = world.contains("vary_glass_ior") ? world.at("vary_glass_ior").asBool() : false;

  std::string meshPath = joinPath(folder, j.at("mesh").asString());
  std::vector<ObjShape> shapes;
  std::vector<RawMaterial> raw;
  parseObj(meshPath, folder, sc.vertices, sc.normals, sc.texcoords, shapes, raw);

  // Scene.cpp:65-118.  Texture names are relative to the scene folder; a
  // file used by several materials in the same role is loaded once.
  std::map<std::string, int> loaded;
  auto texture = [&](const std::string& name, uint32_t channels, bool srgb) -> int {
    if (name.empty()) return -1;
    const std::string path = joinPath(folder, name);
    const std::string key = path + (channels == 1 ? "|1" : srgb ? "|srgb" : "|3");
    auto it = loaded.find(key);
    if (it != loaded.end()) return it->second;
    sc.textures.push_back(loadTexture(path, channels, srgb));
    return loaded[key] = (int)sc.textures.size() - 1;
  };
  std::vector<uint8_t> tinted;
  for (const RawMaterial& m : raw) {
    Material out;
    out.texNormal = texture(m.bumpTex, 3, false);
    out.texAlbedo = texture(m.diffuseTex, 3, true);
    out.texRoughness = texture(m.specularHighlightTex, 1, false);
    out.texMetallic = texture(m.metallicTex, 1, false);
    out.name = m.name;
    for (int i = 0; i < 3; ++i) out.albedo[i] = m.diffuse[i];
    out.transmission = 1.0f - m.dissolve;
    out.varyIOR = sc.varyGlassIOR && out.transmission > 0.0f;
    float nSpecular = m.shininess;
    out.roughness = 1.0f - hm::fsqrt(nSpecular / 1000.0f);
    out.metallic = m.ambient[0];
    out.specular = m.specular[0];
    out.ior = m.ior;
    out.anisotropy = m.anisotropy;
    for (int i = 0; i < 3; ++i) out.emission[i] = m.emission[i];
    sc.materials.push_back(out);
    tinted.push_back(m.diffuse[0] > 0.0f || m.diffuse[1] > 0.0f || m.diffuse[2] > 0.0f || m.specular[0] > 0.0f ||
                     m.specular[1] > 0.0f || m.specular[2] > 0.0f);
  }
  buildTriangles(sc, shapes, tinted);
}

// RenderJob.cpp:21-28
void loadJob(const std::string& renderJobPath, Job& job) {
  job.jobPath = renderJobPath;
  JsonValue j = readJsonFile(renderJobPath);
  if (j.contains("settings")) settingsFromJson(j.at("settings"), job.settings);
  std::string folder = parentDir(renderJobPath);
  loadScene(joinPath(folder, j.at("scene").asString()), job.scene);
}

}  // namespace hasty

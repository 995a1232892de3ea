// Device-side building blocks of the wavefront path tracer (gfx950).
//
// Everything a path vertex needs -- closest-hit traversal of the 4-wide BVH,
// the surface interaction, PrincipledBRDF sample/evaluate/pdf, camera rays,
// light and environment sampling, the media stack -- written per lane for
// 64-wide wavefronts.  Values that feed a branch decision (ray origins and
// directions, lobe probabilities, rejection tests, pdf comparisons) are
// computed with the same IEEE operation order as the reference, so a path
// takes the same decisions on the GPU as on the CPU; accumulated radiance
// may differ by float reassociation only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hmath.h"
#include "rng.h"

namespace hdev {

constexpr double kPi = 3.1415926535897932384;          // core/src/VMath.cpp:7
constexpr double kInvPi = 1.0 / 3.1415926535897932384;  // core/src/VMath.cpp:8
constexpr float kFltEps = 1.1920928955078125e-07f;
constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr int kMaxMedia = 8;

// ------------------------------------------------------------ vectors ----
struct V3 {
  float x, y, z;
};
__host__ __device__ inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ inline V3 operator*(float s, V3 a) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ inline V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ inline V3 mul(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline float norm(V3 a) { return hm::fsqrt(dot(a, a)); }
__device__ inline V3 normalize(V3 a) { return a / norm(a); }
__device__ inline bool isZero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ inline bool isOnes(V3 a) { return a.x == 1.0f && a.y == 1.0f && a.z == 1.0f; }
__device__ inline float sq(float f) { return f * f; }
__device__ inline float mixf(float a, float b, float l) { return (1.0f - l) * a + l * b; }
__device__ inline V3 mix3(V3 a, V3 b, float l) {
  return V3{(1.0f - l) * a.x + l * b.x, (1.0f - l) * a.y + l * b.y, (1.0f - l) * a.z + l * b.z};
}
__device__ inline float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ inline V3 load3(const float* p) { return V3{p[0], p[1], p[2]}; }
// Loads from scene arrays (device allocations) through global-address-space
// pointers: a pointer the compiler cannot trace to a kernel argument (a
// batched job's scene) would otherwise be read with flat loads, which also
// wait on the LDS counter.  Only for global memory, never for locals.
typedef float HNf4 __attribute__((ext_vector_type(4)));
template <class T>
__device__ inline T ldg(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
__device__ inline float4 ldg4(const float4* p) {
  const HNf4 v = *(const __attribute__((address_space(1))) HNf4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline V3 load3g(const float* p) { return V3{ldg(p), ldg(p + 1), ldg(p + 2)}; }
// Copy of a plain struct in global memory, by words.
template <class T>
__device__ inline T ldgStruct(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "word-sized");
  const uint32_t* src = (const uint32_t*)p;
  uint32_t w[sizeof(T) / 4];
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = ldg(src + i);
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}
__device__ inline V3 f4xyz(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ inline float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// VMath.h:461-488
__device__ inline V3 orthonormalized(V3 v1, V3 v2) {
  V3 r = v2 - v1 * dot(v2, v1);
  float len = norm(r);
  return r / len;
}
__device__ inline V3 anyOrthonormal(V3 v1) {
  if (__builtin_fabsf(v1.x) < 0.7f) return orthonormalized(v1, v3(1.0f, 0.0f, 0.0f));
  return orthonormalized(v1, v3(0.0f, 1.0f, 0.0f));
}
__device__ inline V3 orthonormalizedOrAny(V3 v1, V3 v2) {
  V3 r = v2 - v1 * dot(v2, v1);
  float len = norm(r);
  if (len < 1e-7f) return anyOrthonormal(v1);
  return r / len;
}
// Rodrigues rotation taking `from` onto `to` (VMath.h:558-602).
struct Rot {
  float a;
  V3 b, c, d;
};
__device__ inline Rot makeRot(V3 from, V3 to) {
  Rot r;
  float cosA = dot(from, to);
  float onePlus = 1.0f + cosA;
  if (__builtin_fabsf(onePlus) > kFltEps) {
    V3 k = cross(from, to);
    r.a = cosA;
    r.b = k;
    r.c = k;
    r.d = k / onePlus;
  } else {
    V3 k = anyOrthonormal(from);
    r.a = -1.0f;
    r.b = v3(0.0f, 0.0f, 0.0f);
    r.c = k;
    r.d = k * 2.0f;
  }
  return r;
}
__device__ inline V3 applyRot(const Rot& r, V3 x) { return x * r.a + cross(r.b, x) + r.c * dot(r.d, x); }
__device__ inline V3 reflectAcross(V3 wo, V3 n) {  // VMath.h:652-668
  V3 inc = -wo;
  return inc - n * (2.0f * dot(n, inc));
}
template <int N>
__device__ inline float powci(float f) {  // VMath.h:29-46
  if constexpr (N == 0) {
    return 1.0f;
  } else {
    float q = powci<N / 2>(f);
    if constexpr (N % 2 == 1) return f * (q * q);
    else return q * q;
  }
}

// ---------------------------------------------------------- scene data ----
struct DevMaterial {
  float albedo[3];
  float roughness, alpha, metallic, specular, ior, anisotropy, transmission;
  float emission[3];
  int varyIOR, isLight, hasDiffuseLobe;
  // image textures, indices into DevScene::texs (-1 = the constant above):
  // albedo, roughness, metallic, normal map; `textured` = any of them
  int32_t tex[4];
  int textured;
};
// Texels of one image texture: 1 or 3 floats each, row-major, row 0 = top.
struct DevTexture {
  const float* px;
  uint32_t w, h;
};

struct DevScene {
  const float4* nodes;  // 8 float4 per 4-wide node (or node half)
  int bvhWidth;         // children per node: 4 or 8
  const float4* tris;   // 3 float4 per leaf triangle
  const float* triPos;  // 9 per global triangle
  const float* triNrm;
  const int32_t* triMat;
  const uint32_t* triGeom;
  const uint32_t* shapeTriOffset;  // first global triangle of each shape (primID = tri - offset)
  const uint8_t* triSmooth;
  const DevMaterial* mats;
  const DevTexture* texs;
  const float* triUV;        // 6 per triangle (texture scenes only)
  const uint8_t* triHasUV;   // (texture scenes only)
  int texturedMedia;         // some material has an albedo texture
  const uint32_t* lightTris;
  const double* lightCdf;
  uint32_t numLights;
  float totalLightArea;
  const float* envRGB;
  const float* envPdf;
  const double* envCdf;
  const uint32_t* envCdfGuide;  // kEnvGuide + 1 search brackets
  uint32_t envW, envH;
  float envValueSum;
  int hasEnv;
  float bg[3];
  int hasLight, hasSurfaceLight;
  float camPos[3], camFwd[3], camRight[3], camUp[3];
  float fovDegree, aperture, focalDistance, bladeRotation;
  int numBlades;
  // per-scene constants of the camera, computed on the host with the same
  // hmath functions: tan(fov / 2) and the n-gon's edge normal (cos, sin of
  // half a sector)
  float tanHalfFov, ngonNx, ngonNy;
  int maxDepth, rrStart;
  float rrQ;
  int useMIS;
  uint32_t width, height, seed;
};

// ----------------------------------------------------------- traversal ----
struct RayPre {
  V3 o, d;
  int kx, ky, kz;
  float Sx, Sy, Sz;
  // box tests: t of plane L along an axis = (L - o) * inv.  Closest-hit
  // traversal takes min / max of each axis' two planes with inv = 1 / d.
  // Occlusion traversal picks the near and far plane up front (fewer
  // instructions, one more address per plane) with binv = 1 / d whose
  // components are clamped to |d| >= 1e-20 so that no t is NaN; near* = byte
  // offset in a node of the float4 holding each axis' near planes (lo when
  // binv >= 0, where (lo - o) binv <= (hi - o) binv).
  V3 inv, binv;
  uint32_t nearX, nearY, nearZ;
};
__device__ inline float c3(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ inline RayPre prepareRay(V3 o, V3 d) {
  RayPre r;
  r.o = o;
  r.d = d;
  r.inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  r.binv = v3(1.0f / copysignf(fmaxf(__builtin_fabsf(d.x), 1e-20f), d.x),
              1.0f / copysignf(fmaxf(__builtin_fabsf(d.y), 1e-20f), d.y),
              1.0f / copysignf(fmaxf(__builtin_fabsf(d.z), 1e-20f), d.z));
  r.nearX = r.binv.x >= 0.0f ? 0u : 16u;
  r.nearY = r.binv.y >= 0.0f ? 32u : 48u;
  r.nearZ = r.binv.z >= 0.0f ? 64u : 80u;
  float ax = __builtin_fabsf(d.x), ay = __builtin_fabsf(d.y), az = __builtin_fabsf(d.z);
  r.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  r.kx = r.kz + 1 == 3 ? 0 : r.kz + 1;
  r.ky = r.kx + 1 == 3 ? 0 : r.kx + 1;
  if (c3(d, r.kz) < 0.0f) {
    int t = r.kx;
    r.kx = r.ky;
    r.ky = t;
  }
  float dz = c3(d, r.kz);
  r.Sx = c3(d, r.kx) / dz;
  r.Sy = c3(d, r.ky) / dz;
  r.Sz = 1.0f / dz;
  return r;
}

struct Hit {
  float t, u, v;
  int32_t tri;  // global triangle index, -1 = miss
};

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), t > tnear.
__device__ inline bool intersectTri(const RayPre& r, float4 p0, float4 p1, float4 p2, float tnear, float* tOut,
                                    float* uOut, float* vOut) {
  V3 A = f4xyz(p0) - r.o, B = f4xyz(p1) - r.o, C = f4xyz(p2) - r.o;
  float Akz = c3(A, r.kz), Bkz = c3(B, r.kz), Ckz = c3(C, r.kz);
  float Ax = c3(A, r.kx) - r.Sx * Akz;
  float Ay = c3(A, r.ky) - r.Sy * Akz;
  float Bx = c3(B, r.kx) - r.Sx * Bkz;
  float By = c3(B, r.ky) - r.Sy * Bkz;
  float Cx = c3(C, r.kx) - r.Sx * Ckz;
  float Cy = c3(C, r.ky) - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
    U = (float)(CxBy - CyBx);
    double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
    V = (float)(AxCy - AyCx);
    double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
    W = (float)(BxAy - ByAx);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float Az = r.Sz * Akz;
  float Bz = r.Sz * Bkz;
  float Cz = r.Sz * Ckz;
  float T = U * Az + V * Bz + W * Cz;
  float t = T / det;
  if (!(t > tnear)) return false;
  *tOut = t;
  *uOut = V / det;
  *vOut = W / det;
  return true;
}

// Ize's conservative bound: tmax scaled by 1 + 2*gamma(3).
constexpr float kTmaxScale = 1.0f + 2.0f * 3.0f * 0x1p-24f / (1.0f - 3.0f * 0x1p-24f);
constexpr float kCullSlack = 1.001f;

// Register compare-exchange of (entry distance, child ref) pairs.
__device__ inline void cswap(float& ta, uint32_t& ra, float& tb, uint32_t& rb) {
  bool sw = tb < ta;
  float t = sw ? tb : ta;
  uint32_t r = sw ? rb : ra;
  tb = sw ? ta : tb;
  rb = sw ? ra : rb;
  ta = t;
  ra = r;
}

// Shared BVH walk.  CLOSEST: smallest t, ties to the smaller triangle index.
// !CLOSEST (occlusion): stops at the first triangle that beats the bound
// (t < tBound, or t == tBound with index < idBound) and reports it.
// The stack keeps (ref, entry distance): the first LDS_DEPTH entries in LDS,
// interleaved by lane (entry i of a lane at lds[i * laneStride]), deeper ones
// in the lane's global overflow slice.  Children are ordered near-to-far with
// a 4-input sorting network held in registers.
//
// The walk is a state (Walk) advanced one step at a time (walkStep: one
// inner node, or one leaf and the pop after it).  Written this way the
// compiled loop is shorter than the single-loop form with its break flag
// (closest-hit launches -5.6 %, occlusion -8 % on the GGX grid).
struct Walk {
  RayPre r;
  float tBound;
  int32_t idBound;
  Hit best;
  uint32_t ref;  // current node (the root is always inner)
  int sp;        // stack depth
  uint32_t nNodes, nTris;
};
__device__ inline void walkInit(Walk& w, const RayPre& r, float tBound, int32_t idBound) {
  w.r = r;
  w.tBound = tBound;
  w.idBound = idBound;
  w.best.t = tBound;
  w.best.u = w.best.v = 0.0f;
  w.best.tri = -1;
  w.ref = 0;
  w.sp = 0;
}
// One step; true when the walk has ended (w.best then holds the result,
// see walkHit).
// Entry distances of the four children of one 128-byte node half (a 4-wide
// node, or half of an 8-wide one): +inf where the child is not entered
// before tb.  The exit is scaled by 1 + 6u (Ize, "Robust BVH ray
// traversal") so that rounding never culls a box the exact ray enters.  An
// empty slot's planes are all +inf, so its entry is +inf or its exit -inf:
// the closest-hit test needs no branch on the child ref (the occlusion test
// keeps it: faster there).
template <bool CLOSEST>
__device__ inline void testHalf(const DevScene& s, const RayPre& r, uint32_t half, float tb, float& e0, float& e1,
                                float& e2, float& e3, uint32_t& r0, uint32_t& r1, uint32_t& r2, uint32_t& r3) {
  int4 refs;
  if constexpr (CLOSEST) {
    const float4* n = s.nodes + (size_t)half * 8;
    const float4 lox = ldg4(n), hix = ldg4(n + 1), loy = ldg4(n + 2), hiy = ldg4(n + 3), loz = ldg4(n + 4),
                 hiz = ldg4(n + 5), rf = ldg4(n + 6);
    refs = make_int4(__float_as_int(rf.x), __float_as_int(rf.y), __float_as_int(rf.z), __float_as_int(rf.w));
#define HDEV_SLAB(LX, HX, LY, HY, LZ, HZ, RF, TOUT)                                          \
  {                                                                                          \
    float t0x = (LX - r.o.x) * r.inv.x, t1x = (HX - r.o.x) * r.inv.x;                        \
    float t0y = (LY - r.o.y) * r.inv.y, t1y = (HY - r.o.y) * r.inv.y;                        \
    float t0z = (LZ - r.o.z) * r.inv.z, t1z = (HZ - r.o.z) * r.inv.z;                        \
    float tmin = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));           \
    float tmax = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z)) * kTmaxScale; \
    bool h = (tmin <= tmax) & (tmax > 0.0f) & (tmin <= tb);                                      \
    TOUT = h ? tmin : __builtin_inff();                                                      \
  }
    HDEV_SLAB(lox.x, hix.x, loy.x, hiy.x, loz.x, hiz.x, refs.x, e0)
    HDEV_SLAB(lox.y, hix.y, loy.y, hiy.y, loz.y, hiz.y, refs.y, e1)
    HDEV_SLAB(lox.z, hix.z, loy.z, hiy.z, loz.z, hiz.z, refs.z, e2)
    HDEV_SLAB(lox.w, hix.w, loy.w, hiy.w, loz.w, hiz.w, refs.w, e3)
#undef HDEV_SLAB
  } else {
    // 32-bit byte offsets from the node array (< 4 GiB of nodes)
    const char* nb = reinterpret_cast<const char*>(s.nodes);
    const uint32_t no = half * 128u;
    auto plane = [&](uint32_t off) { return ldg4(reinterpret_cast<const float4*>(nb + (no + off))); };
    const float4 nx = plane(r.nearX), fx = plane(r.nearX ^ 16u), ny = plane(r.nearY), fy = plane(r.nearY ^ 16u),
                 nz = plane(r.nearZ), fz = plane(r.nearZ ^ 16u), rf = plane(96u);
    refs = make_int4(__float_as_int(rf.x), __float_as_int(rf.y), __float_as_int(rf.z), __float_as_int(rf.w));
#define HDEV_SLAB(NX, FX, NY, FY, NZ, FZ, RF, TOUT)                                                \
  {                                                                                                \
    float tmin = fmaxf(fmaxf((NX - r.o.x) * r.binv.x, (NY - r.o.y) * r.binv.y), (NZ - r.o.z) * r.binv.z);  \
    float tmax = fminf(fminf((FX - r.o.x) * r.binv.x, (FY - r.o.y) * r.binv.y), (FZ - r.o.z) * r.binv.z) * \
                 kTmaxScale;                                                                       \
    bool h = (uint32_t)(RF) != kLeafFlag && tmin <= tmax && tmax > 0.0f && tmin <= tb;            \
    TOUT = h ? tmin : __builtin_inff();                                                            \
  }
    HDEV_SLAB(nx.x, fx.x, ny.x, fy.x, nz.x, fz.x, refs.x, e0)
    HDEV_SLAB(nx.y, fx.y, ny.y, fy.y, nz.y, fz.y, refs.y, e1)
    HDEV_SLAB(nx.z, fx.z, ny.z, fy.z, nz.z, fz.z, refs.z, e2)
    HDEV_SLAB(nx.w, fx.w, ny.w, fy.w, nz.w, fz.w, refs.w, e3)
#undef HDEV_SLAB
  }
  r0 = (uint32_t)refs.x, r1 = (uint32_t)refs.y, r2 = (uint32_t)refs.z, r3 = (uint32_t)refs.w;
}

// One step; true when the walk has ended (w.best then holds the result,
// see walkHit).  W: children per node (4 or 8; an 8-wide node is two
// 128-byte halves).
template <int LDS_DEPTH, int OVF_DEPTH, bool CLOSEST, int W = 4>
__device__ inline bool walkStep(const DevScene& s, Walk& w, uint2* ldsGeneric, int laneStride, uint2* ovf) {
  static_assert(W == 4 || W == 8, "4- or 8-wide nodes");
  // The stack's LDS part through an LDS-qualified pointer: the compiler
  // then keeps LDS and overflow accesses apart (ds_* and global_*) instead
  // of merging them into one flat access, which would wait on both counters.
  // Entries are (ref, entry distance) packed as ref | bits(t) << 32.
  typedef __attribute__((address_space(3))) unsigned long long LdsEntry;
  LdsEntry* lds = (LdsEntry*)ldsGeneric;
  const RayPre& r = w.r;
  // push a child (farthest first); the BVH depth is checked at load so that
  // (W - 1) * depth + 1 entries fit, the guard only keeps every write inside
  // this lane's slice
#define HDEV_PUSH(RR, EE)                                                                \
  if ((EE) != __builtin_inff() && w.sp < LDS_DEPTH + OVF_DEPTH) {                       \
    uint2 e_ = make_uint2(RR, __float_as_uint(EE));                                     \
    if (w.sp < LDS_DEPTH) lds[w.sp * laneStride] = ((unsigned long long)e_.y << 32) | e_.x; \
    else ovf[w.sp - LDS_DEPTH] = e_;                                                    \
    ++w.sp;                                                                             \
  }
  if ((w.ref & kLeafFlag) == 0) {
    w.nNodes++;
    const float tb = (CLOSEST && w.best.tri < 0) ? __builtin_inff() : w.best.t * kCullSlack;
    if constexpr (W == 4) {
      float e0, e1, e2, e3;
      uint32_t r0, r1, r2, r3;
      testHalf<CLOSEST>(s, r, w.ref, tb, e0, e1, e2, e3, r0, r1, r2, r3);
      // children entered have finite entry distances; after the sort they
      // come first; continue with the nearest
      if (fminf(fminf(e0, e1), fminf(e2, e3)) != __builtin_inff()) {
        cswap(e0, r0, e1, r1);
        cswap(e2, r2, e3, r3);
        cswap(e0, r0, e2, r2);
        cswap(e1, r1, e3, r3);
        cswap(e1, r1, e2, r2);
        HDEV_PUSH(r3, e3)
        HDEV_PUSH(r2, e2)
        HDEV_PUSH(r1, e1)
        w.ref = r0;
        return false;
      }
    } else {
      float e0, e1, e2, e3, e4, e5, e6, e7;
      uint32_t r0, r1, r2, r3, r4, r5, r6, r7;
      testHalf<CLOSEST>(s, r, 2u * w.ref, tb, e0, e1, e2, e3, r0, r1, r2, r3);
      testHalf<CLOSEST>(s, r, 2u * w.ref + 1u, tb, e4, e5, e6, e7, r4, r5, r6, r7);
      if (fminf(fminf(fminf(e0, e1), fminf(e2, e3)), fminf(fminf(e4, e5), fminf(e6, e7))) != __builtin_inff()) {
        // Batcher's odd-even merge sort of 8 (19 compare-exchanges)
        cswap(e0, r0, e1, r1);
        cswap(e2, r2, e3, r3);
        cswap(e4, r4, e5, r5);
        cswap(e6, r6, e7, r7);
        cswap(e0, r0, e2, r2);
        cswap(e1, r1, e3, r3);
        cswap(e4, r4, e6, r6);
        cswap(e5, r5, e7, r7);
        cswap(e1, r1, e2, r2);
        cswap(e5, r5, e6, r6);
        cswap(e0, r0, e4, r4);
        cswap(e1, r1, e5, r5);
        cswap(e2, r2, e6, r6);
        cswap(e3, r3, e7, r7);
        cswap(e2, r2, e4, r4);
        cswap(e3, r3, e5, r5);
        cswap(e1, r1, e2, r2);
        cswap(e3, r3, e4, r4);
        cswap(e5, r5, e6, r6);
        HDEV_PUSH(r7, e7)
        HDEV_PUSH(r6, e6)
        HDEV_PUSH(r5, e5)
        HDEV_PUSH(r4, e4)
        HDEV_PUSH(r3, e3)
        HDEV_PUSH(r2, e2)
        HDEV_PUSH(r1, e1)
        w.ref = r0;
        return false;
      }
    }
  } else if (w.ref != kLeafFlag) {
    uint32_t count = (w.ref >> 24) & 0x7Fu;
    uint32_t first = w.ref & 0x00FFFFFFu;
    for (uint32_t k = 0; k < count; ++k) {
      const float4* tr = s.tris + (size_t)(first + k) * 3;
      float4 p0 = ldg4(tr), p1 = ldg4(tr + 1), p2 = ldg4(tr + 2);
      int32_t id = (int32_t)__float_as_uint(p0.w);
      w.nTris++;
      float t, u, v;
      if (intersectTri(r, p0, p1, p2, 1e-6f, &t, &u, &v)) {
        if (CLOSEST) {
          if (w.best.tri < 0 || t < w.best.t || (t == w.best.t && id < w.best.tri)) {
            w.best.t = t;
            w.best.u = u;
            w.best.v = v;
            w.best.tri = id;
          }
        } else if (t < w.tBound || (t == w.tBound && id < w.idBound)) {
          w.best.t = t;
          w.best.u = u;
          w.best.v = v;
          w.best.tri = id;
          return true;
        }
      }
    }
  }
#undef HDEV_PUSH
  // pop the nearest remaining entry that can still hold a closer hit
  while (w.sp > 0) {
    --w.sp;
    uint2 e;
    if (w.sp < LDS_DEPTH) {
      const unsigned long long v = lds[w.sp * laneStride];
      e = make_uint2((uint32_t)v, (uint32_t)(v >> 32));
    } else {  // rare; volatile so that it is not merged with the LDS load
      const volatile uint2* o = ovf + (w.sp - LDS_DEPTH);
      e = make_uint2(o->x, o->y);
    }
    float lim = (CLOSEST && w.best.tri < 0) ? __builtin_inff() : w.best.t * kCullSlack;
    if (__uint_as_float(e.y) <= lim) {
      w.ref = e.x;
      return false;
    }
  }
  return true;
}
// The result of an ended walk (a closest-hit miss reports t = +inf).
template <bool CLOSEST>
__device__ inline Hit walkHit(const Walk& w) {
  Hit h = w.best;
  if (CLOSEST && h.tri < 0) h.t = __builtin_inff();
  return h;
}

template <int LDS_DEPTH, int OVF_DEPTH, bool CLOSEST, int W = 4>
__device__ inline Hit traverseW(const DevScene& s, const RayPre& r, float tBound, int32_t idBound, uint2* ldsGeneric,
                                int laneStride, uint2* ovf, uint32_t* nodeCount, uint32_t* triCount) {
  Walk w;
  walkInit(w, r, tBound, idBound);
  w.nNodes = w.nTris = 0;
  while (!walkStep<LDS_DEPTH, OVF_DEPTH, CLOSEST, W>(s, w, ldsGeneric, laneStride, ovf)) {
  }
  if (nodeCount) *nodeCount += w.nNodes;
  if (triCount) *triCount += w.nTris;
  return walkHit<CLOSEST>(w);
}
// W = 0: the scene's node width (s.bvhWidth, the same for every lane).
template <int LDS_DEPTH, int OVF_DEPTH, bool CLOSEST, int W = 0>
__device__ inline Hit traverse(const DevScene& s, const RayPre& r, float tBound, int32_t idBound, uint2* ldsGeneric,
                               int laneStride, uint2* ovf, uint32_t* nodeCount, uint32_t* triCount) {
  if constexpr (W != 0) {
    return traverseW<LDS_DEPTH, OVF_DEPTH, CLOSEST, W>(s, r, tBound, idBound, ldsGeneric, laneStride, ovf, nodeCount,
                                                      triCount);
  } else {
    if (s.bvhWidth == 8)
      return traverseW<LDS_DEPTH, OVF_DEPTH, CLOSEST, 8>(s, r, tBound, idBound, ldsGeneric, laneStride, ovf,
                                                         nodeCount, triCount);
    return traverseW<LDS_DEPTH, OVF_DEPTH, CLOSEST, 4>(s, r, tBound, idBound, ldsGeneric, laneStride, ovf, nodeCount,
                                                       triCount);
  }
}

template <int LDS_DEPTH, int OVF_DEPTH, int W = 0>
__device__ inline Hit traceClosest(const DevScene& s, V3 o, V3 d, uint2* lds, int laneStride, uint2* ovf,
                                   uint32_t* nodeCount, uint32_t* triCount) {
  RayPre r = prepareRay(o, d);
  return traverse<LDS_DEPTH, OVF_DEPTH, true, W>(s, r, __builtin_inff(), 0, lds, laneStride, ovf, nodeCount,
                                                 triCount);
}

// Is there a triangle whose hit beats (tBound, idBound)?  Returns its index or -1.
template <int LDS_DEPTH, int OVF_DEPTH, int W = 0>
__device__ inline int32_t findBlocker(const DevScene& s, const RayPre& r, float tBound, int32_t idBound, uint2* lds,
                                      int laneStride, uint2* ovf, uint32_t* nodeCount, uint32_t* triCount) {
  Hit h = traverse<LDS_DEPTH, OVF_DEPTH, false, W>(s, r, tBound, idBound, lds, laneStride, ovf, nodeCount, triCount);
  return h.tri;
}

// ---------------------------------------------------------- interaction ----
struct Interaction {
  int32_t tri;
  uint32_t geomID;
  int32_t mat;
  float h;
  V3 x, ng, ns, tangent, bitangent;
};

// Scene::getInteraction (Scene.cpp:202-255) for explicit barycentrics and
// geometric normal.
__device__ inline Interaction makeInteractionB(const DevScene& s, int32_t tri, float b0, float b1, float b2, V3 ng) {
  Interaction it;
  const float* p = s.triPos + (size_t)tri * 9;
  V3 p0 = load3g(p), p1 = load3g(p + 3), p2 = load3g(p + 6);
  it.tri = tri;
  it.geomID = ldg(s.triGeom + tri);
  it.mat = ldg(s.triMat + tri);
  it.ng = ng;
  it.x = p0 * b0 + p1 * b1 + p2 * b2;
  if (ldg(s.triSmooth + tri)) {
    const float* q = s.triNrm + (size_t)tri * 9;
    V3 n0 = load3g(q), n1 = load3g(q + 3), n2 = load3g(q + 6);
    it.ns = normalize(n0 * b0 + n1 * b1 + n2 * b2);
    float h = 0.0f;
    V3 xv[3] = {p0, p1, p2};
    V3 nv[3] = {n0, n1, n2};
    float bb[3] = {b0, b1, b2};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      float nvng = dot(it.ng, nv[i]);
      float hi = nvng == 0.0f ? 1e6f : (dot(xv[i], nv[i]) - dot(it.x, nv[i])) / nvng;
      h += bb[i] * bb[i] * hi;
    }
    it.h = h;
  } else {
    it.ns = it.ng;
    it.h = 0.0f;
  }
  it.tangent = orthonormalizedOrAny(it.ng, v3(-it.x.z, 0.0f, it.x.x));
  it.bitangent = cross(it.ng, it.tangent);
  return it;
}
// Scene::rayHit + getInteraction (Scene.cpp:281-303): barycentrics (1-u-v, u, v),
// geometric normal normalize((v1-v0) x (v2-v0)).
__device__ inline Interaction makeInteraction(const DevScene& s, int32_t tri, float u, float v) {
  const float* p = s.triPos + (size_t)tri * 9;
  V3 p0 = load3g(p), p1 = load3g(p + 3), p2 = load3g(p + 6);
  V3 ng = normalize(cross(p1 - p0, p2 - p0));
  return makeInteractionB(s, tri, 1.0f - u - v, u, v, ng);
}
// Scene.cpp:257-270
__device__ inline V3 offsetOrigin(const Interaction& it, bool outside) {
  float eps = 1e-5f;
  float hmin = fmaxf(eps, norm(it.x) * eps);
  float k = 1.01f * (outside ? fmaxf(0.0f, it.h) + hmin : fminf(0.0f, it.h) - hmin);
  return it.x + it.ng * k;
}

// ------------------------------------------------------------ sampling ----
// Sample.cpp:77-108
__device__ inline V3 sampleCosHemisphere(hrng::Stream& rng, V3 n, float* pdf) {
  V3 dir;
  float costheta;
  Rot rot = makeRot(v3(0.0f, 0.0f, 1.0f), n);
  do {
    float xi1 = rng.uniform01f();
    float xi2 = rng.uniform01f();
    float phi = 2.0f * (float)kPi * xi2;
    costheta = hm::fsqrt(xi1);
    float sintheta = hm::fsqrt(1.0f - xi1);
    float sinphi, cosphi;
    hm::sincosf(phi, &sinphi, &cosphi);
    V3 rv = v3(sintheta * cosphi, sintheta * sinphi, costheta);
    dir = normalize(applyRot(rot, rv));
    costheta = dot(dir, n);
  } while (costheta <= 0.0f);
  *pdf = (float)kInvPi * costheta;
  return dir;
}
__device__ inline float cosHemispherePdf(V3 n, V3 d) {  // Sample.cpp:109-117
  float c = dot(n, d);
  if (c < 0.0f) return 0.0f;
  return (float)kInvPi * c;
}

// ---------------------------------------------------------------- GGX ----
__device__ inline float DGGX(float nh, float th, float bh, float at, float ab) {  // BRDF.cpp:149-153
  return (float)kInvPi / (at * ab * sq(sq(th / at) + sq(bh / ab) + sq(nh)));
}
__device__ inline float smithGGX(float nv, float tv, float bv, float at, float ab) {  // BRDF.cpp:155-159
  float lambda = 0.5f * (hm::fsqrt(1.0f + (sq(at * tv) + sq(ab * bv)) / sq(nv)) - 1.0f);
  return 1.0f / (1.0f + lambda);
}
// Heitz 2018 visible-normal sample, a pure function of xi (BRDF.cpp:163-185).
__device__ inline V3 sampleVNDF(V3 w, float ax, float ay, float xi1, float xi2) {
  V3 Vh = normalize(v3(ax * w.x, ay * w.y, w.z));
  V3 T1 = cross(Vh, v3(0.0f, 0.0f, 1.0f));
  if (dot(T1, T1) < 1e-12f) T1 = v3(1.0f, 0.0f, 0.0f);
  else T1 = normalize(T1);
  V3 T2 = cross(T1, Vh);
  float r = hm::fsqrt(xi1);
  float a = 1.0f / (1.0f + Vh.z);
  float phi = (float)kPi * ((xi2 < a) ? (xi2 / a) : (1.0f + (xi2 - a) / (1.0f - a)));
  float sphi, cphi;
  hm::sincosf(phi, &sphi, &cphi);
  float t1 = r * cphi;
  float t2 = r * sphi * (xi2 < a ? 1.0f : Vh.z);
  float t3 = hm::fsqrt(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2));
  V3 Nh = T1 * t1 + T2 * t2 + Vh * t3;
  return normalize(v3(ax * Nh.x, ay * Nh.y, fmaxf(0.0f, Nh.z)));
}
// BRDF.cpp:186-213
__device__ inline V3 sampleVNDFGlobal(hrng::Stream& rng, V3 n, float at, float ab, V3 t, V3 b, V3 w1, float* pdf) {
  if (dot(w1, n) <= 0.0f) {
    *pdf = 0.0f;
    return -n;
  }
  V3 Ne;
  do {
    V3 Ve = v3(dot(t, w1), dot(b, w1), dot(n, w1));
    float xi1 = rng.uniform01f();
    float xi2 = rng.uniform01f();
    V3 Nl = sampleVNDF(Ve, at, ab, xi1, xi2);
    Ne = normalize(t * Nl.x + b * Nl.y + n * Nl.z);
  } while (dot(Ne, w1) <= 0.0f);
  V3 w2 = normalize(reflectAcross(w1, Ne));
  float G1 = smithGGX(dot(n, w1), dot(t, w1), dot(b, w1), at, ab);
  float D = DGGX(dot(n, Ne), dot(t, Ne), dot(b, Ne), at, ab);
  float pH = G1 * dot(w1, Ne) * D / dot(w1, n);
  *pdf = pH / (1e-7f + 4.0f * __builtin_fabsf(dot(w1, Ne)));
  return w2;
}
// BRDF.cpp:214-226
__device__ inline float vndfGlobalPdf(V3 n, float at, float ab, V3 t, V3 b, V3 w1, V3 w2) {
  if (dot(w1, n) <= 0.0f) return 0.0f;
  V3 Ne = normalize(w1 + w2);
  float G1 = smithGGX(dot(n, w1), dot(t, w1), dot(b, w1), at, ab);
  float D = DGGX(dot(n, Ne), dot(t, Ne), dot(b, Ne), at, ab);
  float pH = G1 * fmaxf(0.0f, dot(w1, Ne)) * D / dot(w1, n);
  return pH / (1e-7f + 4.0f * __builtin_fabsf(dot(w1, Ne)));
}

// -------------------------------------------------------- Principled ----
__device__ inline float fresnelDielectric(float ci, float ct, float ni, float nt) {  // BRDF.cpp:74-81
  float rp = (nt * ci - ni * ct) / (nt * ci + ni * ct);
  float ro = (ni * ci - nt * ct) / (ni * ci + nt * ct);
  return 0.5f * (rp * rp + ro * ro);
}
struct Prob {
  bool noStrategy, refractionPossible;
  float pD, pS, pR, FD, cosT;
};
// BRDF.cpp:264-307 (flags: 1 diffuse, 2 concentrated)
__device__ inline Prob computeProbability(const DevMaterial& m, float cos_o, float ior_o, float ior_refr_o,
                                          int flags) {
  Prob r;
  r.cosT = 0.0f;
  float relIOR = ior_o / ior_refr_o;
  float sin_iSq = fmaxf(0.0f, 1.0f - cos_o * cos_o);
  float sin_tSq = (relIOR * relIOR) * sin_iSq;
  r.refractionPossible = !(sin_tSq >= 1.0f);
  if (r.refractionPossible) {
    r.cosT = hm::fsqrt(1.0f - sin_tSq);
    r.FD = fresnelDielectric(cos_o, r.cosT, ior_o, ior_refr_o);
  } else {
    r.FD = 1.0f;
  }
  float mD = (1.0f - m.transmission) * (1.0f - m.metallic) * (1.0f - 0.5f * m.specular);
  float mS = mixf(mixf(0.5f * m.specular, r.FD, m.transmission), 1.0f, m.metallic);
  float mR = (float)((1.0 - (double)r.FD) * (double)m.transmission * (1.0 - (double)m.metallic));
  if (!(flags & 1)) mD = 0.0f;
  if (!(flags & 2)) {
    mS = 0.0f;
    mR = 0.0f;
  }
  float sum = mD + mS + mR;
  r.noStrategy = sum == 0.0f;
  if (r.noStrategy) {
    r.pD = r.pS = r.pR = 0.0f;
  } else {
    r.pD = mD / sum;
    r.pS = mS / sum;
    r.pR = mR / sum;
  }
  return r;
}
// BRDF.cpp:604-628 (no normal maps)
__device__ inline V3 shadingNormal(const Interaction& it, V3 wo) {
  float sideNg = dot(it.ng, wo) < 0.0f ? -1.0f : 1.0f;
  V3 ns = it.ns;
  float sideNs = dot(ns, wo) < 0.0f ? -1.0f : 1.0f;
  if (sideNg != sideNs) ns = it.ng;
  return ns * sideNg;
}
__device__ inline float normalScale(V3 wi, V3 ns, V3 ng) {  // BRDF.cpp:239-242
  return __builtin_fabsf(dot(wi, ns)) / (1e-7f + __builtin_fabsf(dot(wi, ng)));
}
struct Frame {
  float at, ab;
  V3 t, b;
};
__device__ inline Frame anisoFrame(const Interaction& it, V3 ns, float alpha, float aniso) {  // BRDF.cpp:308-316
  Frame f;
  float aspect = hm::fsqrt(1.0f - 0.9f * aniso);
  f.at = alpha / aspect;
  f.ab = alpha * aspect;
  Rot rot = makeRot(it.ng, ns);
  f.t = applyRot(rot, it.tangent);
  f.b = applyRot(rot, it.bitangent);
  return f;
}
__device__ inline V3 albedoOf(const DevMaterial& m) { return v3(m.albedo[0], m.albedo[1], m.albedo[2]); }
// BRDF.cpp:319-326
__device__ inline V3 evalDiffuse(float dotNsWi, const DevMaterial& m, float nScale) {
  if (dotNsWi < 0.0f) return v3(0.0f, 0.0f, 0.0f);
  float k = (float)kInvPi * nScale * (1.0f - 0.5f * m.specular);
  float s = (float)(((double)k * (1.0 - (double)m.transmission)) * (1.0 - (double)m.metallic));
  return albedoOf(m) * s;
}
// BRDF.cpp:328-394 (transmissive materials are smooth: rough glass is rejected at load)
__device__ inline V3 evalConcentrated(float dotNsWi, const DevMaterial& m, V3 wo, V3 wi, const Frame& f, V3 ns,
                                      float nScale, bool forceHitSpecular, float FD) {
  if (dotNsWi < 0.0f) return v3(0.0f, 0.0f, 0.0f);
  V3 h = normalize(wo + wi);
  float lh = __builtin_fabsf(dot(wi, h));
  float c5 = powci<5>(1.0f - lh);
  float nh = dot(ns, h);
  float nv = __builtin_fabsf(dot(ns, wo));
  float nl = __builtin_fabsf(dot(ns, wi));
  float Go = smithGGX(nv, dot(f.t, wo), dot(f.b, wo), f.at, f.ab);
  float Gi = smithGGX(nl, dot(f.t, wi), dot(f.b, wi), f.at, f.ab);
  float G = Go * Gi;
  float D;
  if (m.alpha == 0.0f) D = forceHitSpecular ? 4.0f * nl : 0.0f;
  else D = DGGX(nh, dot(f.t, h), dot(f.b, h), f.at, f.ab);
  float specDiel = 0.5f * m.specular;
  V3 specTrans = v3(0.0f, 0.0f, 0.0f);
  V3 alb = albedoOf(m);
  V3 specMetal = v3(alb.x * (1.0f - c5) + c5, alb.y * (1.0f - c5) + c5, alb.z * (1.0f - c5) + c5);
  if (m.transmission > 0.0f) specTrans = forceHitSpecular ? v3(FD, FD, FD) : v3(0.0f, 0.0f, 0.0f);
  float k = D * G * nScale / __builtin_fabsf(1e-7f + 4.0f * nv * nl);
  return mix3(mix3(v3(specDiel, specDiel, specDiel), specTrans, m.transmission), specMetal, m.metallic) * k;
}
__device__ inline float iorOf(const DevMaterial& m, int wavelength) {  // BRDF.cpp:589-603
  if (!m.varyIOR) return m.ior;
  if (wavelength == 0) return m.ior * (1.43f / 1.45f);
  if (wavelength == 1) return m.ior * (1.45f / 1.45f);
  if (wavelength == 2) return m.ior * (1.47f / 1.45f);
  return __builtin_nanf("");
}
struct BsdfEval {
  V3 fD, fC;
};
// BRDF.cpp:630-679 (both lobes)
__device__ inline BsdfEval bsdfEvaluate(const DevMaterial& m, const Interaction& it, V3 wo, V3 wi) {
  BsdfEval e;
  V3 ns = shadingNormal(it, wo);
  float dotNsWi = dot(ns, wi);
  float nScale = normalScale(wi, ns, it.ng);
  e.fD = evalDiffuse(dotNsWi, m, nScale);
  Frame f = anisoFrame(it, ns, m.alpha, m.anisotropy);
  e.fC = evalConcentrated(dotNsWi, m, wo, wi, f, ns, nScale, false, __builtin_nanf(""));
  return e;
}
// BRDF.cpp:702-742
__device__ inline float bsdfPdf(const DevMaterial& m, const Interaction& it, V3 wo, V3 wi, float outsideIOR) {
  float dotNgWo = dot(it.ng, wo);
  V3 ns = shadingNormal(it, wo);
  float dotNsWo = dot(ns, wo);
  bool fromOutside = dotNgWo >= 0.0f;
  float ior_o = fromOutside ? outsideIOR : m.ior;
  float ior_t = fromOutside ? m.ior : outsideIOR;
  Prob pr = computeProbability(m, dotNsWo, ior_o, ior_t, 3);
  if (m.alpha == 0.0f) return pr.pD * cosHemispherePdf(ns, wi);
  Frame f = anisoFrame(it, ns, m.alpha, m.anisotropy);
  float pdfSpec = vndfGlobalPdf(ns, f.at, f.ab, f.t, f.b, wo, wi);
  float pdfDiff = cosHemispherePdf(ns, wi);
  return pr.pS * pdfSpec + pr.pD * pdfDiff;
}
struct BsdfSample {
  bool outside;
  V3 dir, tD, tC;
  float pdf;
};
// BRDF.cpp:397-579 + member wrapper 681-700
__device__ inline BsdfSample bsdfSample(const DevMaterial& m, hrng::Stream& rng, const Interaction& it, V3 wo,
                                        int wavelength, float outsideIOR, bool adjoint, int flags) {
  BsdfSample res;
  V3 ns = shadingNormal(it, wo);
  float iorMat = iorOf(m, wavelength);
  float dotNgWo = dot(it.ng, wo);
  float dotNsWo = dot(ns, wo);
  res.outside = dotNgWo >= 0.0f;
  float ior_o = res.outside ? outsideIOR : iorMat;
  float ior_r = res.outside ? iorMat : outsideIOR;
  Prob pr = computeProbability(m, dotNsWo, ior_o, ior_r, flags);
  if (pr.noStrategy) {
    res.dir = v3(1.0f, 0.0f, 0.0f);
    res.tD = res.tC = v3(0.0f, 0.0f, 0.0f);
    res.pdf = 0.0f;
    return res;
  }
  Frame f = anisoFrame(it, ns, m.alpha, m.anisotropy);
  V3 fD = v3(0.0f, 0.0f, 0.0f), fC = v3(0.0f, 0.0f, 0.0f);
  float xi = rng.uniform01f();
  int strategy = xi < pr.pR ? 3 : (xi < (pr.pR + pr.pS) ? 2 : 1);
  if (strategy == 2) {
    if (m.alpha == 0.0f) {
      res.dir = reflectAcross(wo, ns);
      res.pdf = 1e12f;
      float dotNsWi = dot(ns, res.dir);
      float nScale = normalScale(res.dir, ns, it.ng);
      fC = evalConcentrated(dotNsWi, m, wo, res.dir, f, ns, nScale, true, pr.FD) * (1e12f / pr.pS);
    } else {
      float pdfSpec;
      res.dir = sampleVNDFGlobal(rng, ns, f.at, f.ab, f.t, f.b, wo, &pdfSpec);
      float pdfDiff = cosHemispherePdf(ns, res.dir);
      res.pdf = pr.pS * pdfSpec + pr.pD * pdfDiff;
      float dotNsWi = dot(ns, res.dir);
      float nScale = normalScale(res.dir, ns, it.ng);
      fD = evalDiffuse(dotNsWi, m, nScale);
      fC = evalConcentrated(dotNsWi, m, wo, res.dir, f, ns, nScale, false, pr.FD);
    }
  } else if (strategy == 1) {
    float pdfDiff;
    res.dir = sampleCosHemisphere(rng, ns, &pdfDiff);
    float dotNsWi = dot(ns, res.dir);
    float nScale = normalScale(res.dir, ns, it.ng);
    if (m.alpha == 0.0f) {
      res.pdf = pr.pD * pdfDiff;
      fD = evalDiffuse(dotNsWi, m, nScale);
    } else {
      float pdfSpec = vndfGlobalPdf(ns, f.at, f.ab, f.t, f.b, wo, res.dir);
      res.pdf = pr.pS * pdfSpec + (1.0f - pr.pS) * pdfDiff;
      fD = evalDiffuse(dotNsWi, m, nScale);
      fC = evalConcentrated(dotNsWi, m, wo, res.dir, f, ns, nScale, false, pr.FD);
    }
  } else {
    res.outside = !res.outside;
    res.pdf = 1e12f;
    if (!pr.refractionPossible) {
      res.dir = v3(1.0f, 0.0f, 0.0f);
    } else {
      float r = ior_o / ior_r;
      res.dir = normalize(wo * (-r) + ns * (r * dotNsWo - pr.cosT));
      float scale = adjoint ? 1.0f : sq(r);
      float fc = res.pdf / pr.pR * scale * (1.0f - pr.FD) / __builtin_fabsf(dot(res.dir, it.ng));
      V3 mask = wavelength == -1 ? v3(1.0f, 1.0f, 1.0f)
                                 : v3(wavelength == 0 ? 1.0f : 0.0f, wavelength == 1 ? 1.0f : 0.0f,
                                      wavelength == 2 ? 1.0f : 0.0f);
      fC = mask * fc;
    }
  }
  float mm = __builtin_fabsf(dot(res.dir, it.ng)) / res.pdf;
  res.tD = fD * mm;
  res.tC = fC * mm;
  return res;
}

// ------------------------------------------------------------ textures ----
__device__ inline int floorInt32(double x) {  // VMath.h:18-22
  int i = (int)x;
  return i - (i > x);
}
constexpr uint32_t kTexelNoUV = 0xFFFFFFFFu;
// Texture::evaluateUV with Repeat addressing in u and v (Texture.h:119-140):
// the nearest texel's index.  Two guards where the reference's result is
// undefined: a coordinate that is not finite or not below 2^31 in magnitude
// reads as 0, and a coordinate that rounds to 1 after wrapping reads the last
// column / row.
__device__ inline uint32_t texelOf(const DevTexture& t, float u, float v) {
  if (!(__builtin_fabsf(u) < 2147483648.0f)) u = 0.0f;
  if (!(__builtin_fabsf(v) < 2147483648.0f)) v = 0.0f;
  float uu = u - (float)floorInt32(u);
  float vv = 1.0f - v;
  float om = vv - (float)floorInt32(vv);
  int col = (int)(uu * (float)t.w);
  int row = (int)(om * (float)t.h);
  col = col < 0 ? 0 : (col > (int)t.w - 1 ? (int)t.w - 1 : col);
  row = row < 0 ? 0 : (row > (int)t.h - 1 ? (int)t.h - 1 : row);
  return (uint32_t)row * t.w + (uint32_t)col;
}
__device__ inline float alphaOf(float roughness) {  // PrincipledBRDF::computeAlpha, BRDF.cpp:253-263
  float a = roughness * roughness;
  if (a < 1e-2f && a != 0.0f) a = 1e-2f;
  return a;
}
// The textured parameters of a hit (Texture.cpp:9-18 evaluated by
// PrincipledBRDF, BRDF.cpp:604-750): albedo, roughness (with alpha),
// metallic (with the diffuse-lobe flag) from the texels under the hit's uv =
// uv0 b0 + uv1 b1 + uv2 b2 (Scene.cpp:243-248), and the normal map's normal
// normalize(t x + b y + ng z) with (x, y, z) = 2 texel - 1 in place of the
// default shading normal (getShadingNormal then applies its side test).  A
// triangle without uv reads the NoUV values (1, 0, 0.7) and -1
// (Texture.h:84-101).  *albedoTexel = the albedo texel (kTexelNoUV without
// uv), which the media stack keeps for the entered surface.
__device__ inline void applyTextures(const DevScene& s, DevMaterial& m, Interaction& it, float b0, float b1,
                                     float b2, uint32_t* albedoTexel) {
  const bool hasUV = ldg(s.triHasUV + it.tri) != 0;
  float u = 0.0f, v = 0.0f;
  if (hasUV) {
    const float* q = s.triUV + (size_t)it.tri * 6;
    u = ldg(q) * b0 + ldg(q + 2) * b1 + ldg(q + 4) * b2;
    v = ldg(q + 1) * b0 + ldg(q + 3) * b1 + ldg(q + 5) * b2;
  }
  if (m.tex[0] >= 0) {
    const DevTexture t = ldgStruct(s.texs + m.tex[0]);
    const uint32_t k = hasUV ? texelOf(t, u, v) : kTexelNoUV;
    const V3 a = hasUV ? load3g(t.px + (size_t)k * 3) : v3(1.0f, 0.0f, 0.7f);
    m.albedo[0] = a.x, m.albedo[1] = a.y, m.albedo[2] = a.z;
    *albedoTexel = k;
  }
  if (m.tex[1] >= 0) {
    const DevTexture t = ldgStruct(s.texs + m.tex[1]);
    m.roughness = hasUV ? ldg(t.px + texelOf(t, u, v)) : -1.0f;
    m.alpha = alphaOf(m.roughness);
  }
  if (m.tex[2] >= 0) {
    const DevTexture t = ldgStruct(s.texs + m.tex[2]);
    m.metallic = hasUV ? ldg(t.px + texelOf(t, u, v)) : -1.0f;
    m.hasDiffuseLobe = ((1.0f - 0.5f * m.specular) * (1.0 - (double)m.transmission) * (1.0 - (double)m.metallic) > 0.0)
                           ? 1 : 0;
  }
  if (m.tex[3] >= 0) {
    const DevTexture t = ldgStruct(s.texs + m.tex[3]);
    const V3 c = hasUV ? load3g(t.px + (size_t)texelOf(t, u, v) * 3) : v3(1.0f, 0.0f, 0.7f);
    const V3 tn = v3(2.0f * c.x - 1.0f, 2.0f * c.y - 1.0f, 2.0f * c.z - 1.0f);
    it.ns = normalize(it.tangent * tn.x + it.bitangent * tn.y + it.ng * tn.z);
  }
}

// ---------------------------------------------------------- media stack ----
// LightRayInfo (LightRayInfo.cpp:8-125), entries = geomID << 16 | material.
// An entry's transmittance is the albedo of the surface where the path
// entered it (LightRayInfo.cpp:14).  For a material with an albedo texture
// that is a texel: scenes with such materials keep each entry's texel index
// in `txs` (kMaxMedia per path slot in HBM), and tx0 holds the first entry's.
struct Media {
  uint32_t e[kMaxMedia];
  int n;
  bool overflow;
  uint32_t tx0;
};
__device__ inline V3 mediaEntryAlbedo(const DevScene& s, uint32_t e, uint32_t texel) {
  const DevMaterial* m = s.mats + (e & 0xFFFFu);
  const int32_t ta = ldg(&m->tex[0]);
  if (ta < 0) return albedoOf(*m);
  if (texel == kTexelNoUV) return v3(1.0f, 0.0f, 0.7f);
  return load3g(ldg(&s.texs[ta].px) + (size_t)texel * 3);
}
__device__ inline V3 mediaTransmittance(const DevScene& s, const Media& md) {
  if (md.n == 0) return v3(1.0f, 1.0f, 1.0f);
  return mediaEntryAlbedo(s, md.e[0], md.tx0);
}
// Index of the first entry of shape geomID, or -1.
__device__ inline int mediaFind(const Media& md, uint32_t geomID) {
  int idx = -1;
#pragma unroll
  for (int i = 0; i < kMaxMedia; ++i)
    if (idx < 0 && i < md.n && (md.e[i] >> 16) == geomID) idx = i;
  return idx;
}
// updateMedia; `texel` is the hit's albedo texel and `txs` the path's texel
// slots (nullptr when no material has an albedo texture).
__device__ inline void mediaUpdate(Media& md, const Interaction& it, V3 wFrom, V3 wTo, uint32_t texel,
                                   uint32_t* txs) {
  float ndFrom = dot(it.ng, wFrom);
  float ndTo = dot(it.ng, wTo);
  // (entries are addressed with constant indices only, so the stack stays
  // in registers)
  if (ndFrom >= 0.0f && ndTo < 0.0f) {
    if (md.n < kMaxMedia) {
      const uint32_t v = (it.geomID << 16) | (uint32_t)it.mat;
#pragma unroll
      for (int i = 0; i < kMaxMedia; ++i)
        if (i == md.n) md.e[i] = v;
      if (txs) txs[md.n] = texel;
      if (md.n == 0) md.tx0 = texel;
      md.n += 1;
    } else {
      md.overflow = true;
    }
  } else if (ndFrom < 0.0f && ndTo >= 0.0f) {
    const int idx = mediaFind(md, it.geomID);
    if (idx != -1) {
#pragma unroll
      for (int j = 0; j < kMaxMedia - 1; ++j)
        if (j >= idx && j < md.n - 1) md.e[j] = md.e[j + 1];
      if (txs) {
        for (int j = idx; j < md.n - 1; ++j) txs[j] = txs[j + 1];
        if (idx == 0 && md.n > 1) md.tx0 = txs[0];
      }
      md.n -= 1;
    }
  }
}
// The transmittance the stack would have after mediaUpdate(md, it, wFrom,
// wTo, texel, ...), without changing it (the light ray's copy of the state).
__device__ inline V3 mediaTransmittanceAfter(const DevScene& s, const Media& md, const Interaction& it, V3 wFrom,
                                             V3 wTo, uint32_t texel, const uint32_t* txs) {
  float ndFrom = dot(it.ng, wFrom);
  float ndTo = dot(it.ng, wTo);
  if (ndFrom >= 0.0f && ndTo < 0.0f) {
    if (md.n == 0) return mediaEntryAlbedo(s, (it.geomID << 16) | (uint32_t)it.mat, texel);
  } else if (ndFrom < 0.0f && ndTo >= 0.0f) {
    if (mediaFind(md, it.geomID) == 0) {
      if (md.n == 1) return v3(1.0f, 1.0f, 1.0f);
      return mediaEntryAlbedo(s, md.e[1], txs ? txs[1] : 0u);
    }
  }
  return mediaTransmittance(s, md);
}
__device__ inline float mediaOutsideIOR(const DevScene& s, const Media& md, const Interaction& it, V3 wFrom,
                                        int wavelength) {
  if (dot(wFrom, it.ng) < 0.0f) {
    const int idx = mediaFind(md, it.geomID);
    if (idx == -1) return 1.0f;
    if (md.n == 1) return 1.0f;
    uint32_t other = idx == 0 ? md.e[1] : md.e[0];
    return iorOf(s.mats[other & 0xFFFFu], wavelength);
  }
  if (md.n == 0) return 1.0f;
  return iorOf(s.mats[md.e[0] & 0xFFFFu], wavelength);
}
// PathTracing.h:46-59
__device__ inline V3 beersLaw(V3 T, float d) {
  if (isOnes(T)) return v3(1.0f, 1.0f, 1.0f);
  return v3(T.x <= 0.0f ? 0.0f : hm::powf_pos(T.x, d), T.y <= 0.0f ? 0.0f : hm::powf_pos(T.y, d),
            T.z <= 0.0f ? 0.0f : hm::powf_pos(T.z, d));
}

// ------------------------------------------------------- environment ----
__device__ inline void sphericalFromDir(V3 d, float* theta, float* phi) {  // VMath.h:692-699
  *theta = hm::acosf(d.y);
  float p = hm::atan2f(d.z, d.x);
  *phi = p < 0.0f ? p + (float)(2.0 * kPi) : p;
}
// Texture::evaluateUV with U repeat / V clamp (Texture.h:119-140)
__device__ inline V3 envLookup(const DevScene& s, float u, float v) {
  if (u != u || v != v) return v3(0.0f, 0.0f, 0.0f);  // direction off the unit sphere
  float uu = u - (float)floorInt32(u);
  float om = fminf(fmaxf(1.0f - v, 0.0f), 1.0f - kFltEps);
  int col = (int)(uu * s.envW);
  int row = (int)(om * s.envH);
  if (col > (int)s.envW - 1) col = (int)s.envW - 1;
  return load3g(s.envRGB + ((size_t)row * s.envW + col) * 3);
}
__device__ inline V3 evalEnvironment(const DevScene& s, V3 dir) {  // Scene.cpp:482-485
  if (!s.hasEnv) return v3(s.bg[0], s.bg[1], s.bg[2]);
  float th, ph;
  sphericalFromDir(dir, &th, &ph);
  return envLookup(s, ph * (float)(0.5 * kInvPi), 1.0f - th * (float)kInvPi);
}
// ImageDistribution::evalDensity / EnvironmentTexture::evalSamplingDensity
__device__ inline float envPdf(const DevScene& s, V3 dir) {
  float th, ph;
  sphericalFromDir(dir, &th, &ph);
  float u = ph * (float)(0.5 * kInvPi), v = 1.0f - th * (float)kInvPi;
  float sinT = hm::sinf(th);
  float W = (float)s.envW, H = (float)s.envH;
  if (u != u || v != v) return 0.0f / fmaxf(1e-7f, 2.0f * (float)(kPi * kPi) * sinT);
  float fx = u * s.envW, fy = (1.0f - v) * s.envH;
  uint32_t x = fx >= (float)(s.envW - 1) ? s.envW - 1 : (uint32_t)fx;
  uint32_t y = fy >= (float)(s.envH - 1) ? s.envH - 1 : (uint32_t)fy;
  float d = W * H * s.envPdf[(size_t)y * s.envW + x] / s.envValueSum;
  return d / fmaxf(1e-7f, 2.0f * (float)(kPi * kPi) * sinT);
}
// upper_bound over the environment CDF, narrowed by a guide table:
// guide[g] = upper_bound(cdf, g / kEnvGuide), so for x in [g, g + 1) / kEnvGuide
// the answer lies in [guide[g], guide[g + 1]] (upper_bound is monotone) and
// the result is exactly the full search's.
constexpr uint32_t kEnvGuide = 65536;
__device__ inline uint32_t envCdfSearch(const DevScene& s, double x) {
  uint32_t n = s.envW * s.envH;
  uint32_t g = (uint32_t)(x * (double)kEnvGuide);
  if (g > kEnvGuide - 1) g = kEnvGuide - 1;
  uint32_t lo = s.envCdfGuide[g], hi = s.envCdfGuide[g + 1];
  uint32_t count = hi - lo;
  while (count > 0) {
    uint32_t step = count >> 1, mid = lo + step;
    if (!(x < s.envCdf[mid])) {
      lo = mid + 1;
      count -= step + 1;
    } else {
      count = step;
    }
  }
  return lo < n ? lo : n - 1;
}
// EnvironmentTexture::sample (Environment.h:59-78, 141-153)
__device__ inline V3 envSample(const DevScene& s, hrng::Stream& rng, float* density) {
  float sinT;
  V3 dir;
  float dens;
  do {
    uint32_t idx = envCdfSearch(s, rng.canonical_double());
    uint32_t W = s.envW, H = s.envH;
    uint32_t x = idx % W, y = idx / W;
    float heightF = (float)H, widthF = (float)W;
    float xi0 = rng.uniform01f();
    float xi1 = rng.uniform01f();
    float left = x / widthF;
    float right = (x + 1.0f) / widthF;
    float lower = (H - (y + 1.0f)) / heightF;
    float upper = (float)(H - y) / heightF;
    float u = (1.0f - xi0) * left + xi0 * right;
    float v = (1.0f - xi1) * lower + xi1 * upper;
    float pxy = s.envPdf[idx] / s.envValueSum;
    dens = widthF * heightF * pxy;
    float phi = u * 2.0f * (float)kPi;
    float theta = (1.0f - v) * (float)kPi;
    float cosT, sinp, cosp;
    hm::sincosf(theta, &sinT, &cosT);
    hm::sincosf(phi, &sinp, &cosp);
    dir = v3(sinT * cosp, cosT, sinT * sinp);
  } while (sinT == 0.0f);
  *density = dens / (2.0f * (float)kPi * (float)kPi * sinT);
  return dir;
}

// --------------------------------------------------------------- lights ----
// Scene.cpp:587-616: light-strategy solid-angle density of (ray, hit).
__device__ inline float lightDensity(const DevScene& s, V3 rayO, V3 rayD, int32_t hitTri, V3 hitX, V3 hitNg) {
  bool surfaceLight = hitTri >= 0 && ldg(&s.mats[ldg(s.triMat + hitTri)].isLight);
  bool infinite = hitTri < 0;
  if (!surfaceLight && !infinite) return 0.0f;
  float pSurf = !s.hasEnv ? 1.0f : (s.hasSurfaceLight ? 0.5f : 0.0f);
  float res;
  if (surfaceLight) {
    float pdfA = 1.0f / s.totalLightArea;
    V3 df = rayO - hitX;
    float diffSq = fmaxf(1e-5f, dot(df, df));
    float pdfW = pdfA * diffSq / (1e-7f + __builtin_fabsf(dot(hitNg, rayD)));
    res = pSurf * pdfW;
    if (s.hasEnv) res += (1.0f - pSurf) * envPdf(s, rayD);
  } else {
    res = s.hasEnv ? (1.0f - pSurf) * envPdf(s, rayD) : 0.0f;
  }
  return res;
}
// Hit point and normalized geometric normal only (for light densities).
__device__ inline void hitPointNormal(const DevScene& s, int32_t tri, float u, float v, V3* x, V3* ng) {
  const float* p = s.triPos + (size_t)tri * 9;
  V3 p0 = load3g(p), p1 = load3g(p + 3), p2 = load3g(p + 6);
  float b0 = 1.0f - u - v;
  *ng = normalize(cross(p1 - p0, p2 - p0));
  *x = p0 * b0 + p1 * u + p2 * v;
}
__device__ inline uint32_t lightCdfSearch(const DevScene& s, double x) {
  uint32_t n = s.numLights;
  uint32_t i = hrng::upper_bound(s.lightCdf, n, x);
  return i < n ? i : n - 1;
}

// ----------------------------------------------------------- camera ----
// Sample.cpp:5-42
__host__ __device__ inline float ngonSector(int numBlades) { return (float)(2.0 * kPi / numBlades); }
__host__ __device__ inline float tanHalfFov(float fovDegree) {
  float fovRad = (float)kPi * fovDegree / 180.0f;
  return hm::tanf(fovRad * 0.5f);
}
__device__ inline void sampleNGon(hrng::Stream& rng, int numBlades, float bladeRotation, float nx, float ny, float* ox,
                                  float* oy) {
  float anglePerSector = ngonSector(numBlades);
  float d = -nx;
  float sx, sy;
  bool outside;
  do {
    float alpha = rng.uniform01f() * (float)(2.0 * kPi);
    float r0 = hm::fsqrt(rng.uniform01f());
    float sa, ca;
    hm::sincosf(alpha, &sa, &ca);
    sx = r0 * ca;
    sy = r0 * sa;
    float r = hm::fsqrt(sx * sx + sy * sy);
    float angle = hm::atan2f(sy, sx);
    angle -= bladeRotation;
    angle += (float)(2.0 * kPi) * (1.0f + kFltEps);
    angle = hm::fmodf(angle, anglePerSector);
    float sang, cang;
    hm::sincosf(angle, &sang, &cang);
    float px = r * cang, py = r * sang;
    outside = (nx * px + ny * py + d) > 0.0f;
  } while (outside);
  *ox = sx;
  *oy = sy;
}
// Camera::computeRay (Camera.cpp:41-81)
__device__ inline void cameraRay(const DevScene& s, hrng::Stream& rng, float px, float py, V3* o, V3* dOut) {
  V3 pos = load3(s.camPos), fwd = load3(s.camFwd), right = load3(s.camRight), up = load3(s.camUp);
  float W = (float)s.width, H = (float)s.height;
  float q = s.tanHalfFov;
  float lx = (px - 0.5f * W) / W * 2.0f;
  float ly = (py - 0.5f * H) / W * 2.0f;
  V3 dir = normalize(fwd + right * (q * lx) - up * (q * ly));
  if (s.aperture > 0.0f) {
    float ox, oy;
    sampleNGon(rng, s.numBlades, s.bladeRotation, s.ngonNx, s.ngonNy, &ox, &oy);
    float t = s.focalDistance * dot(dir, fwd);
    V3 focal = pos + dir * t;
    V3 org = pos + up * (s.aperture * ox) + right * (s.aperture * oy);
    *o = org;
    *dOut = normalize(focal - org);
    return;
  }
  *o = pos;
  *dOut = dir;
}

// Camera::computeRayAngle (Camera.cpp:83-98): angle subtended by one pixel.
__device__ inline float cameraRayAngle(const DevScene& s, float px, float py) {
  float W = (float)s.width, H = (float)s.height;
  float q = s.tanHalfFov;
  float lx1 = (px + 0.5f - 0.5f * W) / W * 2.0f;
  float ly1 = (py + 0.5f - 0.5f * H) / W * 2.0f;
  float lx2 = (px - 0.5f - 0.5f * W) / W * 2.0f;
  float ly2 = (py - 0.5f - 0.5f * H) / W * 2.0f;
  float ax = __builtin_fabsf(hm::atanf(q * lx1) - hm::atanf(q * lx2));
  float ay = __builtin_fabsf(hm::atanf(q * ly1) - hm::atanf(q * ly2));
  return 0.5f * ax + 0.5f * ay;
}

}  // namespace hdev

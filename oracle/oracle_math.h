// ORACLE (test infrastructure only).  CPU restatement of the reference's vector
// math, used by the path-tracing oracle.  Never linked into the product.
//
// Follows core/include/Hasty/VMath.h (Vec2f/Vec3f 50-303, orthonormalisation
// 451-488, RotationBetweenTwoVectors 558-607, reflect 652-668,
// SphericalCoordinates 672-700) operation for operation, with the
// transcendental calls routed through the renderer's deterministic libm
// replacement (hmath.h) so that the oracle and the GPU see identical bits.
#pragma once
#include <math.h>
#include <stdint.h>

#include <cfloat>

#include "../hastyrenderer_amd/csrc/hmath.h"

namespace oracle {

constexpr double Pi = 3.1415926535897932384;          // VMath.cpp:7
constexpr double InvPi = 1.0 / 3.1415926535897932384;  // VMath.cpp:8

inline float square(float f) { return f * f; }  // VMath.h:16
inline int32_t floor_int32(double x) {          // VMath.h:18-22
  int i = (int)x;
  return i - (i > x);
}
inline float fmod1p1(float f) { return f - (float)floor_int32(f); }  // VMath.h:24-27

template <int N>
inline float powci(float f) {  // VMath.h:29-46
  if constexpr (N == 0) {
    return 1.0f;
  } else {
    float q = powci<N / 2>(f);
    if constexpr (N % 2 == 1) return f * (q * q);
    else return q * q;
  }
}

struct Vec2f {
  float v[2];
  Vec2f() : v{0, 0} {}
  Vec2f(float x, float y) : v{x, y} {}
  float& operator[](int i) { return v[i]; }
  const float& operator[](int i) const { return v[i]; }
  Vec2f operator+(const Vec2f& o) const { return Vec2f(v[0] + o[0], v[1] + o[1]); }
  Vec2f operator*(float s) const { return Vec2f(v[0] * s, v[1] * s); }
};
inline float dot(const Vec2f& a, const Vec2f& b) { return a[0] * b[0] + a[1] * b[1]; }
inline float norm(const Vec2f& a) { return hm::fsqrt(dot(a, a)); }

struct Vec3f {
  float v[3];
  Vec3f() : v{0, 0, 0} {}
  Vec3f(float x, float y, float z) : v{x, y, z} {}
  float& operator[](int i) { return v[i]; }
  const float& operator[](int i) const { return v[i]; }
  Vec3f operator+(const Vec3f& o) const { return Vec3f(v[0] + o[0], v[1] + o[1], v[2] + o[2]); }
  Vec3f& operator+=(const Vec3f& o) {
    v[0] += o[0];
    v[1] += o[1];
    v[2] += o[2];
    return *this;
  }
  Vec3f operator/(float s) const { return Vec3f(v[0] / s, v[1] / s, v[2] / s); }
  Vec3f& operator*=(float s) {
    v[0] *= s;
    v[1] *= s;
    v[2] *= s;
    return *this;
  }
  Vec3f& operator/=(float s) {
    v[0] /= s;
    v[1] /= s;
    v[2] /= s;
    return *this;
  }
  Vec3f operator-() const { return Vec3f(-v[0], -v[1], -v[2]); }
  Vec3f operator-(const Vec3f& o) const { return Vec3f(v[0] - o[0], v[1] - o[1], v[2] - o[2]); }
  Vec3f operator*(float s) const { return Vec3f(v[0] * s, v[1] * s, v[2] * s); }
  bool operator==(const Vec3f& o) const { return v[0] == o[0] && v[1] == o[1] && v[2] == o[2]; }
  bool operator!=(const Vec3f& o) const { return !(*this == o); }
  static Vec3f Fill(float f) { return Vec3f(f, f, f); }
  static Vec3f Zero() { return Fill(0.0f); }
  static Vec3f Ones() { return Fill(1.0f); }
};
inline Vec3f operator*(float s, const Vec3f& v) { return v * s; }  // VMath.h:303-306
inline Vec3f cross(const Vec3f& a, const Vec3f& b) {
  return Vec3f(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline Vec3f cwiseProd(const Vec3f& a, const Vec3f& b) { return Vec3f(a[0] * b[0], a[1] * b[1], a[2] * b[2]); }
inline float dot(const Vec3f& a, const Vec3f& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline float normSq(const Vec3f& a) { return dot(a, a); }
inline float norm(const Vec3f& a) { return hm::fsqrt(normSq(a)); }
inline Vec3f normalize(const Vec3f& a) { return a / norm(a); }
inline float sum(const Vec3f& a) { return a[0] + a[1] + a[2]; }
inline float mix(float a, float b, float l) { return (1.0f - l) * a + l * b; }
inline Vec3f mix(const Vec3f& a, const Vec3f& b, float l) {
  Vec3f r;
  for (int i = 0; i < 3; ++i) r[i] = (1.0f - l) * a[i] + l * b[i];
  return r;
}
inline bool isFinite(const Vec3f& a) { return std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]); }

// VMath.h:451-488
inline Vec3f orthonormalized(const Vec3f& v1, const Vec3f& v2) {
  Vec3f r = v2 - dot(v2, v1) * v1;
  float length = norm(r);
  r /= length;
  return r;
}
inline Vec3f anyOrthonormal(const Vec3f& v1) {
  if (std::fabs(v1[0]) < 0.7f) return orthonormalized(v1, Vec3f(1.0f, 0.0f, 0.0f));
  return orthonormalized(v1, Vec3f(0.0f, 1.0f, 0.0f));
}
inline Vec3f orthonormalizedOtherwiseAnyOrthonormal(const Vec3f& v1, const Vec3f& v2) {
  Vec3f r = v2 - dot(v2, v1) * v1;
  float length = norm(r);
  if (length < 1e-7f) r = anyOrthonormal(v1);
  else r /= length;
  return r;
}

// VMath.h:558-607
struct RotationBetweenTwoVectors {
  float a;
  Vec3f b, c, d;
  RotationBetweenTwoVectors(const Vec3f& from, const Vec3f& to) {
    float cosAngle = dot(from, to);
    float onePlusCosAngle = 1.0f + cosAngle;
    if (std::fabs(onePlusCosAngle) > FLT_EPSILON) {
      Vec3f kSinAngle = cross(from, to);
      a = cosAngle;
      b = kSinAngle;
      c = kSinAngle;
      d = kSinAngle / onePlusCosAngle;
    } else {
      Vec3f k = anyOrthonormal(from);
      a = -1.0f;
      b = Vec3f::Zero();
      c = k;
      d = k * 2.0f;
    }
  }
};
inline Vec3f applyRotation(const RotationBetweenTwoVectors& r, const Vec3f& x) {
  return x * r.a + cross(r.b, x) + r.c * dot(r.d, x);
}
inline Vec3f applyRotationInverse(const RotationBetweenTwoVectors& r, const Vec3f& x) {
  return x * r.a - cross(r.b, x) + r.c * dot(r.d, x);
}

struct Ray {
  Vec3f o, d;
  Ray() {}
  Ray(const Vec3f& origin, const Vec3f& direction) : o(origin), d(direction) {}
  const Vec3f& origin() const { return o; }
  const Vec3f& direction() const { return d; }
};

// VMath.h:652-668
inline Vec3f reflect(const Vec3f& incident, const Vec3f& N) { return incident - 2.0f * dot(N, incident) * N; }
inline Vec3f reflectAcross(const Vec3f& outgoing, const Vec3f& N) { return reflect(-outgoing, N); }

// VMath.h:672-700 (theta from +y)
struct ToDirectionResult {
  float cosTheta, sinTheta;
  Vec3f direction;
};
inline ToDirectionResult sphericalToDirection(const Vec2f& c) {
  ToDirectionResult r;
  r.cosTheta = hm::cosf(c[0]);
  r.sinTheta = hm::sinf(c[0]);
  float sinphi = hm::sinf(c[1]);
  float cosphi = hm::cosf(c[1]);
  r.direction = Vec3f(r.sinTheta * cosphi, r.cosTheta, r.sinTheta * sinphi);
  return r;
}
inline Vec2f sphericalFromDirection(const Vec3f& d) {
  float theta = hm::acosf(d[1]);
  float phi = hm::atan2f(d[2], d[0]);
  phi = phi < 0.0f ? phi + float(2.0 * Pi) : phi;
  return Vec2f(theta, phi);
}

}  // namespace oracle

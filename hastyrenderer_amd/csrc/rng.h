// Counter-based random numbers shared by the host and gfx950.
//
// The reference draws from one shishua stream per render thread, consumed
// sequentially over pixels and passes (core/src/Random.cpp:11-47;
// core/src/PathTracing.cpp:102-127), so its images depend on thread
// scheduling and cannot be reproduced pixel by pixel.  This renderer gives
// every path its own stream instead: Philox-4x32-10 (Salmon, Moraes, Dror and
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011) with
//   counter = (stream index, sample index, draw index, domain tag)
//   key     = (seed, 0x48535459)
// and the draw index advanced once per 32-bit word, in the order in which the
// reference consumes its stream along one path (PathTracing.cpp:52, MIS.cpp
// sampling calls, BRDF.cpp:462, Sample.cpp).  The GPU wavefront and the CPU
// oracle therefore draw identical numbers for a path regardless of how the
// paths are scheduled.
//
// uniform01f keeps the reference's mapping float(u & 0x7FFFFF) * 2^-23
// (core/include/Hasty/Random.h:131-137).  Discrete draws replace
// std::discrete_distribution (core/src/Scene.cpp:193, 515;
// core/include/Hasty/Environment.h:56, 62): two words form a double in [0, 1)
// the way libstdc++'s generate_canonical<double> does, and the index is the
// first CDF entry strictly greater than it (upper_bound).
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define HR_HD __host__ __device__ inline
#else
#define HR_HD inline
#endif

namespace hrng {

constexpr uint32_t kKeyTag = 0x48535459u;  // "HSTY"

// Domain tags keep independent uses of one (stream, sample) pair apart.
enum Domain : uint32_t {
  kDomainPath = 1,     // camera + path-tracing vertices
  kDomainPhoton = 2,   // one emitted photon
  kDomainRadius = 3,   // SPPM initial radius probe
  kDomainGuidePhoton = 4,  // a light path training the path guide
};

struct U32x4 {
  uint32_t v[4];
};

HR_HD void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

HR_HD U32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                          uint32_t k1) {
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c0, &hi0, &lo0);
    mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  U32x4 r;
  r.v[0] = c0;
  r.v[1] = c1;
  r.v[2] = c2;
  r.v[3] = c3;
  return r;
}

// One path's stream.  `counter` is the number of words drawn so far.
struct Stream {
  uint32_t seed;
  uint32_t stream;  // pixel index (PT), photon index (SPPM), ...
  uint32_t sample;  // sample / pass index
  uint32_t domain;
  uint32_t counter;

  HR_HD uint32_t next_u32() {
    U32x4 r = philox4x32_10(stream, sample, counter, domain, seed, kKeyTag);
    counter += 1;
    return r.v[0];
  }
  HR_HD float uniform01f() {
    uint32_t u = next_u32();
    return (float)(u & 0x007FFFFFu) * 1.1920928955078125e-07f;
  }
  // libstdc++ generate_canonical<double, 53> over a 32-bit engine: two words,
  // (w0 + w1 * 2^32) / 2^64, clamped below 1.
  HR_HD double canonical_double() {
    uint32_t w0 = next_u32();
    uint32_t w1 = next_u32();
    double r = ((double)w0 + (double)w1 * 4294967296.0) * 5.421010862427522170037e-20;
    if (r >= 1.0) r = 0x1.fffffffffffffp-1;
    return r;
  }
};

HR_HD Stream make_stream(uint32_t seed, uint32_t stream, uint32_t sample, uint32_t domain) {
  Stream s;
  s.seed = seed;
  s.stream = stream;
  s.sample = sample;
  s.domain = domain;
  s.counter = 0;
  return s;
}

// First index i in cdf[0..n) with cdf[i] > x (n if none; callers keep
// cdf[n-1] == 1.0 > x).
HR_HD uint32_t upper_bound(const double* cdf, uint32_t n, double x) {
  uint32_t lo = 0, count = n;
  while (count > 0) {
    uint32_t step = count >> 1;
    uint32_t mid = lo + step;
    if (!(x < cdf[mid])) {
      lo = mid + 1;
      count -= step + 1;
    } else {
      count = step;
    }
  }
  return lo;
}

}  // namespace hrng

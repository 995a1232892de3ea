// Stochastic progressive photon mapping on gfx950: caustic photon map,
// uniform-grid cell table, camera paths with photon gathers and the
// per-pixel progressive statistics (core/src/PhotonMapping.cpp:186-933,
// core/include/Hasty/PhotonMapping.h:12-105, core/include/Hasty/HashCells.h).
//
// One pass:
//   1. cell size: three deterministic reductions over the radius image
//      (mean; eleven growing candidates; ten shrinking candidates);
//   2. photons: batches of photon indices traced in parallel, each thread
//      keeping its stored photons in private slots; a scan of the per-thread
//      counts decides where the sequential emission loop of the reference
//      stops (fewer than storeLimit stored, fewer than emitLimit emitted)
//      and compacts the emitted photons in emission order;
//   3. cells: 63-bit packed cell keys, stable radix sort (key, emission
//      index), cell table (open addressing, key -> first, count);
//   4. camera: one thread per pixel runs the whole MIS path with gathers in
//      the reference's cell order (x, then y, then z; photons of a cell in
//      emission order) and applies the pixel statistics update.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "pt_kernels.h"
#include "sppm.h"

using namespace hdev;

namespace hasty {
namespace {

constexpr int kB = 256;
constexpr unsigned long long kEmptyKey = 0xFFFFFFFFFFFFFFFFull;
constexpr int kKeyBits = 21;
constexpr int kKeyBias = 1 << (kKeyBits - 1);

void hipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define HC(x) hipCheck((x), #x)

__device__ inline uint32_t xcdRemap() {
  uint32_t g = gridDim.x, b = blockIdx.x;
  if ((g & 7u) != 0u) return b;
  return (b & 7u) * (g >> 3) + (b >> 3);
}

// PointHashCells::index (HashCells.h:33-36)
__device__ inline int3 cellOf(V3 p, float cs) {
  return make_int3(floorInt32(p.x / cs), floorInt32(p.y / cs), floorInt32(p.z / cs));
}
__device__ inline bool keyRange(int3 c) {
  return c.x >= -kKeyBias && c.x < kKeyBias && c.y >= -kKeyBias && c.y < kKeyBias && c.z >= -kKeyBias &&
         c.z < kKeyBias;
}
// Lexicographic (x, y, z) order of cells == numeric order of the keys.
__device__ inline unsigned long long packKey(int3 c) {
  return ((unsigned long long)(uint32_t)(c.x + kKeyBias) << (2 * kKeyBits)) |
         ((unsigned long long)(uint32_t)(c.y + kKeyBias) << kKeyBits) | (unsigned long long)(uint32_t)(c.z + kKeyBias);
}
__device__ inline uint32_t hashKey(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return (uint32_t)k;
}

struct SppmDev {
  // pixels (SoA, P entries each)
  float* radius;
  float* N;
  float* Ne;
  float* flux;  // 3P
  float* LSum;  // 3P
  int* Nd;
  float* radii;  // the radius image used for the next pass
  // photons in emission order and in cell order
  float4* phX;   // x, ngwi
  float4* phWi;  // wi
  float4* phF;   // flux
  float4* sX;
  float4* sWi;
  float4* sF;
  unsigned long long* keys;
  uint32_t* idx;
  // cell table
  unsigned long long* tKeys;
  uint2* tVals;
  uint32_t tMask;
  int* bbox;  // min xyz, max xyz of occupied cells
  const float* cellSize;
  uint32_t* counters;
  uint2* ovf;
  // gather queries of the pass's camera paths: record (v, p) of pixel p's
  // v-th diffuse vertex at query[4 (v P + p)]: (x, material), (shading
  // normal, Tb.x), (geometric normal, Tb.y), (W_R, Tb.z)
  float4* query;
  uint32_t* nq;  // queries per pixel
  float* Lpath;  // 3P: LL + LR + LS of the pass's camera path
  // gathers: int64 fixed-point flux with fluxScale = 2^S units per 1
  // (fluxInvScale = 2^-S), and the candidate count above which a wave
  // gathers its queries together
  float fluxScale;
  double fluxInvScale;
  uint32_t coopCandidates;
  float unionFactor;  // union gather when the wave's union box has at most this x the mean cells of its queries
  // gathers of the pass: queries, candidate photons in the cells visited,
  // photons found within the radius
  unsigned long long* gatherStats;
};

enum Counter {
  kCtrStored = 0,
  kCtrEmitted = 1,
  kCtrBatchEmit = 2,
  kCtrKeyOverflow = 3,
  kCtrFluxOverflow = 4,  // gathers whose fixed-point flux sum saturated
  kNumCtr = 5
};

// ------------------------------------------------------------ cell size ----
// PhotonMapping.cpp:522-576 as three reductions.  Each block writes its
// partial sums; a one-block kernel folds them in block order.
constexpr int kCandidates = 11;
__global__ void __launch_bounds__(kB) kCellPartial(const float* radii, uint32_t P, const float* cand, int nCand,
                                                   float* partial) {
  __shared__ float red[kB];
  float acc[kCandidates];
  float cs[kCandidates];
#pragma unroll
  for (int c = 0; c < kCandidates; ++c) {
    acc[c] = 0.0f;
    cs[c] = c < nCand ? cand[c] : 1.0f;
  }
  for (uint32_t p = blockIdx.x * kB + threadIdx.x; p < P; p += gridDim.x * kB) {
    float r = radii[p];
    if (nCand == 0) {
      acc[0] += r;
    } else {
#pragma unroll
      for (int c = 0; c < kCandidates; ++c)
        if (c < nCand) acc[c] += 5.0f * powci<3>(r / cs[c]) + powci<3>(fmaxf(r, cs[c])) * 1e5f;
    }
  }
  int n = nCand == 0 ? 1 : nCand;
#pragma unroll
  for (int c = 0; c < kCandidates; ++c) {
    if (c < n) {  // n is uniform over the block
      red[threadIdx.x] = acc[c];
      __syncthreads();
      for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
      }
      if (threadIdx.x == 0) partial[(size_t)c * gridDim.x + blockIdx.x] = red[0];
      __syncthreads();
    }
  }
}
// stage 0: mean -> 11 growing candidates; stage 1: pick the grown size ->
// 10 shrinking candidates; stage 2: pick the final size.
__global__ void __launch_bounds__(kB) kCellFold(const float* partial, int nBlocks, int stage, uint32_t P, float* cand,
                                                float* state, float* cellSizeOut) {
  __shared__ float red[kB];
  __shared__ float sums[kCandidates];
  int n = stage == 0 ? 1 : (stage == 1 ? 11 : 10);
  for (int c = 0; c < n; ++c) {
    float a = 0.0f;
    for (int b = threadIdx.x; b < nBlocks; b += kB) a += partial[(size_t)c * nBlocks + b];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) sums[c] = red[0];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (stage == 0) {
    float cs = sums[0] / (float)P;
    cand[0] = cs;
    for (int j = 1; j < 11; ++j) cand[j] = cand[j - 1] * 1.5f;
  } else if (stage == 1) {
    // candidates[0] is the mean, [1..10] the growing trials
    float cost = sums[0];
    int a = 0;
    for (int j = 1; j < 11; ++j) {
      if (!(sums[j] < cost)) break;
      cost = sums[j];
      a = j;
    }
    float cs = cand[a];
    state[0] = cs;
    state[1] = cost;
    cand[0] = cs * 0.6666f;
    for (int j = 1; j < 10; ++j) cand[j] = cand[j - 1] * 0.6666f;
  } else {
    float cs = state[0], cost = state[1];
    for (int j = 0; j < 10; ++j) {
      if (!(sums[j] < cost)) break;
      cost = sums[j];
      cs = cand[j];
    }
    *cellSizeOut = cs;
  }
}

// ------------------------------------------------------- initial radii ----
// PhotonMapping.cpp:578-651 (ignoreCosineTerm) times the footprint scale;
// -1 marks pixels without a usable surface (filled on the host).
__global__ void __launch_bounds__(kB) kInitRadii(DevScene s, float footprintScale, float* radii, uint2* ovfAll) {
  __shared__ uint2 stack[kStackLds * kB];
  uint32_t P = s.width * s.height;
  uint32_t gtid = xcdRemap() * kB + threadIdx.x;
  uint2* ovf = ovfAll + (size_t)gtid * kStackOvf;
  for (uint32_t p = gtid; p < P; p += gridDim.x * kB) {
    uint32_t i = p % s.width, j = p / s.width;
    hrng::Stream rng = hrng::make_stream(s.seed, p, 0u, hrng::kDomainRadius);
    float px = (float)i + 0.5f, py = (float)j + 0.5f;
    V3 o, d;
    cameraRay(s, rng, px, py, &o, &d);
    Hit h = traceClosest<kStackLds, kStackOvf>(s, o, d, stack + threadIdx.x, kB, ovf, nullptr, nullptr);
    float result = -1.0f;
    if (h.tri >= 0) {
      float angle = cameraRayAngle(s, px, py);
      Interaction it = makeInteraction(s, h.tri, h.u, h.v);
      float dist = norm(load3(s.camPos) - it.x);
      bool ok = true;
      if (!s.mats[it.mat].hasDiffuseLobe) {
        V3 rd = d;
        for (int k = 0; k < s.maxDepth; ++k) {
          const DevMaterial& m = s.mats[it.mat];
          if (m.hasDiffuseLobe) break;
          BsdfSample bs;
          int tries = 0;
          V3 wo = -rd;
          do {
            // a fresh LightRayInfo each try: no media, wavelength unset
            int wavelength = -1;
            float sp = 1.0f;
            if (m.varyIOR) {
              float p0 = 1.0f / 3.0f;
              float xi = rng.uniform01f();
              wavelength = xi < p0 ? 0 : (xi < p0 + p0 ? 1 : 2);
              sp = p0;
            }
            Media fresh;
            fresh.n = 0;
            fresh.overflow = false;
            fresh.tx0 = 0u;
            float outsideIOR = mediaOutsideIOR(s, fresh, it, wo, wavelength);
            bs = bsdfSample(m, rng, it, wo, wavelength, outsideIOR, false, 2);
            bs.tD = bs.tD / sp;
            bs.tC = bs.tC / sp;
            tries += 1;
          } while (isZero(bs.tC) && tries < 3);
          if (isZero(bs.tC)) break;
          V3 o2 = offsetOrigin(it, bs.outside);
          Hit h2 = traceClosest<kStackLds, kStackOvf>(s, o2, bs.dir, stack + threadIdx.x, kB, ovf, nullptr, nullptr);
          if (h2.tri < 0 || h2.tri == it.tri) {
            ok = false;
            break;
          }
          Interaction it2 = makeInteraction(s, h2.tri, h2.u, h2.v);
          dist += norm(it.x - it2.x);
          it = it2;
          rd = bs.dir;
        }
      }
      if (ok) result = 0.5f * dist * hm::tanf(angle);
    }
    radii[p] = footprintScale * result;
  }
}

// -------------------------------------------------------------- photons ----
// Scene::sampleLightRay (Scene.cpp:527-537) + tracePhoton
// (PhotonMapping.cpp:191-265).  Thread t traces photon index first + t and
// keeps its stores in slots [d * batch + t], d < cap.
__global__ void __launch_bounds__(kB) kTracePhotons(DevScene s, SppmDev g, uint32_t first, uint32_t batch,
                                                    uint32_t emitLimit, uint32_t pass, uint32_t cap, float4* slX,
                                                    float4* slWi, float4* slF, uint32_t* slCount) {
  __shared__ uint2 stack[kStackLds * kB];
  uint32_t gtid = xcdRemap() * kB + threadIdx.x;
  uint2* ovf = g.ovf + (size_t)gtid * kStackOvf;
  const float minQ = 1.0f - hm::powf_pos(0.1f, 1.0f / (float)s.maxDepth);
  for (uint32_t t = gtid; t < batch; t += gridDim.x * kB) {
    uint32_t index = first + t;
    uint32_t stored = 0;
    if (index < emitLimit) {
      hrng::Stream rng = hrng::make_stream(s.seed, index, pass, hrng::kDomainPhoton);
      // sampleSurfaceLightPosition (Scene.cpp:506-526)
      float pPos = 1.0f / s.totalLightArea;
      uint32_t li = lightCdfSearch(s, rng.canonical_double());
      int32_t tri = (int32_t)s.lightTris[li];
      const float* pp = s.triPos + (size_t)tri * 9;
      V3 p0 = load3(pp), p1 = load3(pp + 3), p2 = load3(pp + 6);
      V3 ngL = normalize(cross(p1 - p0, p2 - p0));
      float su = rng.uniform01f();
      float tt = hm::fsqrt(rng.uniform01f());
      float b0 = 1.0f - tt;
      float b2 = tt * su;
      float b1 = fmaxf(0.0f, 1.0f - b0 - b2);
      Interaction itL = makeInteractionB(s, tri, b0, b1, b2, ngL);
      float pDir;
      V3 dir = sampleCosHemisphere(rng, itL.ng, &pDir);
      const DevMaterial& lm = s.mats[itL.mat];
      V3 Le = v3(lm.emission[0], lm.emission[1], lm.emission[2]) * (float)kInvPi;
      V3 flux = Le * (__builtin_fabsf(dot(itL.ng, dir)) / (pPos * pDir));
      V3 ro = offsetOrigin(itL, true), rd = dir;
      Hit h = traceClosest<kStackLds, kStackOvf>(s, ro, rd, stack + threadIdx.x, kB, ovf, nullptr, nullptr);
      if (h.tri >= 0) {
        Media md;
        md.n = 0;
        md.overflow = false;
        md.tx0 = 0u;
        int wavelength = -1;
        Interaction it = makeInteraction(s, h.tri, h.u, h.v);
        // empty media: transmittance is exactly one
        for (int depth = 0; depth < s.maxDepth; ++depth) {
          if (!(__builtin_isfinite(flux.x) && __builtin_isfinite(flux.y) && __builtin_isfinite(flux.z))) break;
          const DevMaterial& m = s.mats[it.mat];
          V3 wo = -rd;
          if (m.hasDiffuseLobe && depth > 0 && stored < cap) {
            size_t sl = (size_t)stored * batch + t;
            slX[sl] = f4(it.x, dot(it.ng, wo));
            slWi[sl] = f4(wo, 0.0f);
            slF[sl] = f4(flux, 0.0f);
            stored++;
          }
          float sp = 1.0f;
          if (wavelength == -1 && m.varyIOR) {
            float q0 = 1.0f / 3.0f;
            float xi = rng.uniform01f();
            wavelength = xi < q0 ? 0 : (xi < q0 + q0 ? 1 : 2);
            sp = q0;
          }
          float outsideIOR = mediaOutsideIOR(s, md, it, wo, wavelength);
          BsdfSample bs = bsdfSample(m, rng, it, wo, wavelength, outsideIOR, true, 2);
          bs.tD = bs.tD / sp;
          bs.tC = bs.tC / sp;
          V3 thrThis = bs.tD + bs.tC;
          if (isZero(thrThis)) break;
          V3 o2 = offsetOrigin(it, bs.outside);
          Hit h2 = traceClosest<kStackLds, kStackOvf>(s, o2, bs.dir, stack + threadIdx.x, kB, ovf, nullptr, nullptr);
          if (h2.tri < 0 || h2.tri == it.tri) break;
          mediaUpdate(md, it, wo, bs.dir, 0u, nullptr);  // photon mapping scenes have no textures
          Interaction it2 = makeInteraction(s, h2.tri, h2.u, h2.v);
          V3 t2 = beersLaw(mediaTransmittance(s, md), norm(it2.x - o2));
          thrThis = mul(thrThis, t2);
          float rq = fmaxf(minQ, 1.0f - norm(thrThis));
          if (rng.uniform01f() <= rq) break;
          flux = mul(flux, thrThis) / (1.0f - rq);
          it = it2;
          rd = bs.dir;
        }
      }
    }
    slCount[t] = stored;
  }
}

// The reference emits photon i while fewer than storeLimit photons are
// stored (PhotonMapping.cpp:861): with exclusive offsets o[t] of this batch,
// photon t is emitted iff stored + o[t] < storeLimit (and i < emitLimit).
// Emitted threads form a prefix; its length goes to kCtrBatchEmit.
__global__ void __launch_bounds__(kB) kEmitCut(SppmDev g, const uint32_t* offs, uint32_t first, uint32_t batch,
                                               uint32_t emitLimit, uint32_t storeLimit) {
  uint32_t t = blockIdx.x * kB + threadIdx.x;
  if (t >= batch) return;
  uint32_t stored = g.counters[kCtrStored];
  bool emitted = first + t < emitLimit && (unsigned long long)stored + offs[t] < storeLimit;
  bool nextEmitted = t + 1 < batch && first + t + 1 < emitLimit &&
                     (unsigned long long)stored + offs[t + 1] < storeLimit;
  if (emitted && !nextEmitted) g.counters[kCtrBatchEmit] = t + 1;
}
__global__ void __launch_bounds__(kB) kEmitCompact(SppmDev g, const uint32_t* offs, const uint32_t* cnt,
                                                   uint32_t batch, const float4* slX, const float4* slWi,
                                                   const float4* slF) {
  uint32_t n = g.counters[kCtrBatchEmit];
  uint32_t base = g.counters[kCtrStored];
  for (uint32_t t = blockIdx.x * kB + threadIdx.x; t < n; t += gridDim.x * kB) {
    uint32_t c = cnt[t], o = base + offs[t];
    for (uint32_t d = 0; d < c; ++d) {
      size_t sl = (size_t)d * batch + t;
      g.phX[o + d] = slX[sl];
      g.phWi[o + d] = slWi[sl];
      g.phF[o + d] = slF[sl];
    }
  }
}
__global__ void kEmitAdvance(SppmDev g, const uint32_t* offs, const uint32_t* cnt) {
  uint32_t n = g.counters[kCtrBatchEmit];
  if (n > 0) g.counters[kCtrStored] += offs[n - 1] + cnt[n - 1];
  g.counters[kCtrEmitted] += n;
  g.counters[kCtrBatchEmit] = 0;
}

// ---------------------------------------------------------------- cells ----
__global__ void __launch_bounds__(kB) kCellKeys(SppmDev g, uint32_t n) {
  __shared__ int box[6];
  if (threadIdx.x < 6) box[threadIdx.x] = threadIdx.x < 3 ? 0x7FFFFFFF : -0x7FFFFFFF;
  __syncthreads();
  float cs = *g.cellSize;
  int3 lo = make_int3(0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF), hi = make_int3(-0x7FFFFFFF, -0x7FFFFFFF, -0x7FFFFFFF);
  for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
    float4 x = g.phX[k];
    int3 c = cellOf(v3(x.x, x.y, x.z), cs);
    if (!keyRange(c)) {
      atomicAdd(&g.counters[kCtrKeyOverflow], 1u);
      c = make_int3(0, 0, 0);
    }
    g.keys[k] = packKey(c);
    g.idx[k] = k;
    lo = make_int3(min(lo.x, c.x), min(lo.y, c.y), min(lo.z, c.z));
    hi = make_int3(max(hi.x, c.x), max(hi.y, c.y), max(hi.z, c.z));
  }
  atomicMin(&box[0], lo.x);
  atomicMin(&box[1], lo.y);
  atomicMin(&box[2], lo.z);
  atomicMax(&box[3], hi.x);
  atomicMax(&box[4], hi.y);
  atomicMax(&box[5], hi.z);
  __syncthreads();
  if (threadIdx.x < 3) atomicMin(&g.bbox[threadIdx.x], box[threadIdx.x]);
  else if (threadIdx.x < 6) atomicMax(&g.bbox[threadIdx.x], box[threadIdx.x]);
}
__global__ void __launch_bounds__(kB) kCellInit(SppmDev g, int clearBox) {
  for (uint32_t k = blockIdx.x * kB + threadIdx.x; k <= g.tMask; k += gridDim.x * kB) g.tKeys[k] = kEmptyKey;
  if (clearBox && blockIdx.x == 0 && threadIdx.x < 6) g.bbox[threadIdx.x] = threadIdx.x < 3 ? 0x7FFFFFFF : -0x7FFFFFFF;
}
// Gather the photon records into cell order and insert each cell's first
// photon into the table; kCellCount then closes each cell at its last photon.
__global__ void __launch_bounds__(kB) kCellBuild(SppmDev g, const unsigned long long* keysSorted,
                                                 const uint32_t* idxSorted, uint32_t n) {
  for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
    uint32_t src = idxSorted[k];
    g.sX[k] = g.phX[src];
    g.sWi[k] = g.phWi[src];
    g.sF[k] = g.phF[src];
    unsigned long long key = keysSorted[k];
    if (k == 0 || keysSorted[k - 1] != key) {
      uint32_t h = hashKey(key) & g.tMask;
      while (true) {
        unsigned long long prev = atomicCAS(&g.tKeys[h], kEmptyKey, key);
        if (prev == kEmptyKey) break;
        h = (h + 1) & g.tMask;
      }
      g.tVals[h] = make_uint2(k, 0u);
    }
  }
}
__global__ void __launch_bounds__(kB) kCellCount(SppmDev g, const unsigned long long* keysSorted, uint32_t n) {
  for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
    unsigned long long key = keysSorted[k];
    if (k + 1 == n || keysSorted[k + 1] != key) {
      uint32_t h = hashKey(key) & g.tMask;
      while (g.tKeys[h] != key) h = (h + 1) & g.tMask;
      g.tVals[h].y = k + 1 - g.tVals[h].x;
    }
  }
}

__device__ inline uint2 cellLookup(const SppmDev& g, unsigned long long key) {
  uint32_t h = hashKey(key) & g.tMask;
  while (true) {
    unsigned long long k = g.tKeys[h];
    if (k == key) return g.tVals[h];
    if (k == kEmptyKey) return make_uint2(0u, 0u);
    h = (h + 1) & g.tMask;
  }
}

// photonMapFluxEstimate (PhotonMapping.cpp:271-296) with
// PointHashCells::radiusNeighbors (HashCells.h:71-98): cells x -> y -> z,
// photons of a cell in emission order, |x - c|^2 <= r^2; the BSDF is the
// diffuse lobe only with outside IOR 1.
//
// When some lane of the wave has many candidate photons (dense maps, large
// radii), the 64 lanes of the wave serve its queries one owner at a time,
// every lane testing a share of the owner's photons (rank k of 64 takes
// photons k, k + 64, ... of each cell, so a cell's run is read in whole
// cache lines), and the lanes' partial sums are added by a butterfly over
// the wave.  The set and the count of photons are the reference's; only the
// FP32 association of the flux sum differs, the same way on every run.
// Otherwise each lane sums its own photons in the reference's order.  In the
// shared gather the owner's cell lookups are spread over the lanes (one hash
// probe each, all in flight together) and the cells' photon runs read back
// from LDS in cell order; `runs` is this wave's 64-entry LDS scratch.
// The sum of a gather's photon terms.  FloatSum: FP32 in the order the
// terms arrive (the reference's, for a lane gathering alone).  FixedSum:
// each term rounded to an integer number of 2^-S units (round to nearest
// even) and summed in int64, so the sum does not depend on the order of the
// terms -- a lane gathering alone, the 64 ranks of a shared gather, any cell
// size, and (for several GPUs) any split of the photons give the same bits.
// A term or a sum outside the int64 range saturates and sets bit 0 of
// `flags` (the pass counts such gathers); a non-finite term sets bit 1 and
// makes the estimate NaN, as the FP32 sum would be.
struct FloatSum {
  V3 v = V3{0.0f, 0.0f, 0.0f};
  __device__ void add(V3 t, const SppmDev&) { v = v + t; }
  __device__ void reduceWave() {
    for (int off = 32; off > 0; off >>= 1) {
      v.x += __shfl_xor(v.x, off);
      v.y += __shfl_xor(v.y, off);
      v.z += __shfl_xor(v.z, off);
    }
  }
  __device__ V3 value(const SppmDev&) const { return v; }
  __device__ bool saturated() const { return false; }
};
struct FixedSum {
  long long x = 0, y = 0, z = 0;
  int flags = 0;
  __device__ static long long toFixed(float t, float scale, int& flags) {
    const float s = t * scale;  // exact: scale is a power of two
    if (!(s == s) || __builtin_isinf(t)) {
      flags |= 2;
      return 0;
    }
    if (!(__builtin_fabsf(s) < 0x1p62f)) {
      flags |= 1;
      return s < 0.0f ? -(1ll << 62) : (1ll << 62);
    }
    return __float2ll_rn(s);
  }
  __device__ static long long addSat(long long a, long long b, int& flags) {
    long long r;
    if (__builtin_add_overflow(a, b, &r)) {
      flags |= 1;
      return a < 0 ? (long long)0x8000000000000000ull : 0x7FFFFFFFFFFFFFFFll;
    }
    return r;
  }
  __device__ void add(V3 t, const SppmDev& g) {
    x = addSat(x, toFixed(t.x, g.fluxScale, flags), flags);
    y = addSat(y, toFixed(t.y, g.fluxScale, flags), flags);
    z = addSat(z, toFixed(t.z, g.fluxScale, flags), flags);
  }
  __device__ static long long shfl64(long long v, int off) {
    const int lo = __shfl_xor((int)(unsigned long long)v, off);
    const int hi = __shfl_xor((int)((unsigned long long)v >> 32), off);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
  }
  __device__ void reduceWave() {
    for (int off = 32; off > 0; off >>= 1) {
      const long long ox = shfl64(x, off), oy = shfl64(y, off), oz = shfl64(z, off);
      flags |= __shfl_xor(flags, off);
      x = addSat(x, ox, flags);
      y = addSat(y, oy, flags);
      z = addSat(z, oz, flags);
    }
  }
  __device__ V3 value(const SppmDev& g) const {
    if (flags & 2) return V3{__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf("")};
    return V3{(float)((double)x * g.fluxInvScale), (float)((double)y * g.fluxInvScale),
              (float)((double)z * g.fluxInvScale)};
  }
  __device__ bool saturated() const { return (flags & 1) != 0; }
};

// photonMapFluxEstimate's term of a photon within the radius, from its
// position, direction and flux records.
template <class Sum>
__device__ inline void photonAdd(const SppmDev& g, float4 x4, float4 wi4, float4 fl4, V3 ns, V3 ng,
                                 const DevMaterial& m, Sum* acc, int* found) {
  const V3 wi = f4xyz(wi4);
  const V3 fl = f4xyz(fl4);
  const V3 fD = evalDiffuse(dot(ns, wi), m, normalScale(wi, ns, ng));
  const float fc = __builtin_fabsf(dot(wi, ng)) / (1e-7f + __builtin_fabsf(x4.w));
  acc->add(mul(fl * fc, fD + v3(0.0f, 0.0f, 0.0f)), g);
  *found += 1;
}
template <class Sum>
__device__ inline void photonTerm(const SppmDev& g, uint32_t q, float4 x4, V3 c, float r2, V3 ns, V3 ng,
                                  const DevMaterial& m, Sum* acc, int* found) {
  const V3 dlt = v3(x4.x, x4.y, x4.z) - c;
  if (!(dot(dlt, dlt) <= r2)) return;
  photonAdd(g, x4, g.sWi[q], g.sF[q], ns, ng, m, acc, found);
}
#ifndef HASTY_PHOTON_LOADS
#define HASTY_PHOTON_LOADS 4
#endif
constexpr uint32_t kPhotonLoads = HASTY_PHOTON_LOADS;
// Photons first, first + stride, ... below end, in that order, with
// kPhotonLoads position loads in flight at a time.
template <class Sum>
__device__ inline void photonRun(const SppmDev& g, uint32_t first, uint32_t end, uint32_t stride, V3 c, float r2,
                                 V3 ns, V3 ng, const DevMaterial& m, Sum* acc, int* found) {
  uint32_t q = first;
  for (; q + (kPhotonLoads - 1u) * stride < end; q += kPhotonLoads * stride) {
    float4 x[kPhotonLoads];
#pragma unroll
    for (uint32_t u = 0; u < kPhotonLoads; ++u) x[u] = g.sX[q + u * stride];
#pragma unroll
    for (uint32_t u = 0; u < kPhotonLoads; ++u) photonTerm(g, q + u * stride, x[u], c, r2, ns, ng, m, acc, found);
  }
  for (; q < end; q += stride) photonTerm(g, q, g.sX[q], c, r2, ns, ng, m, acc, found);
}
// Lanes whose cells hold fewer candidates than g.coopCandidates (default
// 512) gather on their own (a shared gather costs every lane each owner's
// cell lookups).
constexpr uint32_t kCoopCandidatesDefault = 512;
// Union gather: the wave visits the cells of the union of its queries'
// boxes in x -> y -> z order, stages each run of photon positions in LDS 64
// at a time, and every lane tests them against its own query.  A photon
// within a query's radius lies in that query's own box, and the union's
// order restricted to the box is the box's order, so each lane sums exactly
// its photons in the reference's order (the per-lane gather's result), while
// each photon is read once per wave instead of once per query.
template <class Sum>
__device__ inline V3 unionGather(const SppmDev& g, const DevScene& s, bool has, V3 x, V3 ns, V3 ng, int mat,
                                 float r, int3 lo, int3 hi, int3 ulo, int3 uhi, int* M, uint2* runs,
                                 float4* staged) {
  const uint32_t lane = __lane_id();
  Sum est;
  int found = 0;
  DevMaterial m;
  if (has) m = ldgStruct(s.mats + mat);
  const float r2 = r * r;
  const int ny = uhi.y - ulo.y + 1, nz = uhi.z - ulo.z + 1;
  const int nyz = ny * nz, nc = (uhi.x - ulo.x + 1) * nyz;
  for (int base = 0; base < nc; base += 64) {
    const int ci = base + (int)lane;
    uint2 run = make_uint2(0u, 0u);
    if (ci < nc) {
      const int ix = ci / nyz, rem = ci - ix * nyz, iy = rem / nz, iz = rem - iy * nz;
      run = cellLookup(g, packKey(make_int3(ulo.x + ix, ulo.y + iy, ulo.z + iz)));
    }
    runs[lane] = run;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int nb = min(64, nc - base);
    for (int t = 0; t < nb; ++t) {
      const uint2 cell = runs[t];
      if (cell.y == 0u) continue;
      // the cell's coordinates, to skip it on lanes whose box does not hold it
      const int cidx = base + t;
      const int cx = ulo.x + cidx / nyz, crem = cidx % nyz, cy = ulo.y + crem / nz, cz = ulo.z + crem % nz;
      const bool mine = has && cx >= lo.x && cx <= hi.x && cy >= lo.y && cy <= hi.y && cz >= lo.z && cz <= hi.z;
      if (__ballot(mine) == 0ull) continue;
      const uint32_t end = cell.x + cell.y;
      for (uint32_t q0 = cell.x; q0 < end; q0 += 64u) {
        const uint32_t k = q0 + lane;
        staged[lane] = k < end ? g.sX[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t n = min(64u, end - q0);
        if (mine)
          for (uint32_t j = 0; j < n; ++j) photonTerm(g, q0 + j, staged[j], x, r2, ns, ng, m, &est, &found);
        __builtin_amdgcn_wave_barrier();  // staged is rewritten for the next photons
      }
    }
    __builtin_amdgcn_wave_barrier();  // runs are rewritten for the next cells
  }
  if (has && est.saturated()) atomicAdd(&g.counters[kCtrFluxOverflow], 1u);
  *M = found;
  return est.value(g);
}

// The terms of the n <= 64 queued photons fq[head ..] (mod 128), lane l
// taking the l-th; every lane of the wave calls this.
template <class Sum>
__device__ inline void foundTerms(const SppmDev& g, const uint32_t* fq, uint32_t head, uint32_t n, V3 ns, V3 ng,
                                  const DevMaterial& m, Sum* acc, int* found) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint32_t lane = __lane_id();
  if (lane < n) {
    const uint32_t q = fq[(head + lane) & 127u];
    photonAdd(g, g.sX[q], g.sWi[q], g.sF[q], ns, ng, m, acc, found);
  }
  __builtin_amdgcn_wave_barrier();  // the slots are refilled afterwards
}

template <class Sum>
__device__ inline V3 gatherQuery(const SppmDev& g, const DevScene& s, bool has, V3 x, V3 ns, V3 ng, int mat,
                                 float r, int* M, uint2* runs, float4* staged, uint32_t* fq, uint32_t* candidates) {
  const float cs = *g.cellSize;
  {
    // candidates of this lane's own query; if no lane of the wave has many,
    // every lane gathers alone in the reference's order
    uint32_t cand = 0;
    int3 lo = make_int3(0, 0, 0), hi = make_int3(-1, -1, -1);
    if (has) {
      lo = cellOf(x - v3(r, r, r), cs), hi = cellOf(x + v3(r, r, r), cs);
      lo = make_int3(max(lo.x, g.bbox[0]), max(lo.y, g.bbox[1]), max(lo.z, g.bbox[2]));
      hi = make_int3(min(hi.x, g.bbox[3]), min(hi.y, g.bbox[4]), min(hi.z, g.bbox[5]));
      for (int i = lo.x; i <= hi.x; ++i)
        for (int j = lo.y; j <= hi.y; ++j)
          for (int k = lo.z; k <= hi.z; ++k) cand += cellLookup(g, packKey(make_int3(i, j, k))).y;
    }
    *candidates = cand;
    if (__ballot(cand > g.coopCandidates) == 0ull) {
      Sum est;
      int found = 0;
      if (has) {
        const DevMaterial m = ldgStruct(s.mats + mat);
        for (int i = lo.x; i <= hi.x; ++i)
          for (int j = lo.y; j <= hi.y; ++j)
            for (int k = lo.z; k <= hi.z; ++k) {
              const uint2 cell = cellLookup(g, packKey(make_int3(i, j, k)));
              photonRun(g, cell.x, cell.x + cell.y, 1u, x, r * r, ns, ng, m, &est, &found);
            }
        if (est.saturated()) atomicAdd(&g.counters[kCtrFluxOverflow], 1u);
      }
      *M = found;
      return est.value(g);
    }
    // many candidates: the union gather when the queries' boxes overlap
    // enough (the union box holds at most unionFactor x their mean cells)
    if (g.unionFactor > 0.0f) {
      int3 ulo = make_int3(has ? lo.x : 0x7FFFFFFF, has ? lo.y : 0x7FFFFFFF, has ? lo.z : 0x7FFFFFFF);
      int3 uhi = make_int3(has ? hi.x : -0x7FFFFFFF, has ? hi.y : -0x7FFFFFFF, has ? hi.z : -0x7FFFFFFF);
      float own = has ? (float)max(hi.x - lo.x + 1, 0) * (float)max(hi.y - lo.y + 1, 0) *
                            (float)max(hi.z - lo.z + 1, 0)
                      : 0.0f;
      int nHas = has ? 1 : 0;
      for (int off = 32; off > 0; off >>= 1) {
        ulo.x = min(ulo.x, __shfl_xor(ulo.x, off));
        ulo.y = min(ulo.y, __shfl_xor(ulo.y, off));
        ulo.z = min(ulo.z, __shfl_xor(ulo.z, off));
        uhi.x = max(uhi.x, __shfl_xor(uhi.x, off));
        uhi.y = max(uhi.y, __shfl_xor(uhi.y, off));
        uhi.z = max(uhi.z, __shfl_xor(uhi.z, off));
        own += __shfl_xor(own, off);
        nHas += __shfl_xor(nHas, off);
      }
      const float unionCells = nHas > 0 ? (float)(uhi.x - ulo.x + 1) * (float)(uhi.y - ulo.y + 1) *
                                              (float)(uhi.z - ulo.z + 1)
                                        : 0.0f;
      const bool nonEmpty = uhi.x >= ulo.x && uhi.y >= ulo.y && uhi.z >= ulo.z;
      if (nHas > 0 && nonEmpty && unionCells <= g.unionFactor * own / (float)nHas && unionCells < 1.0e6f)
        return unionGather<Sum>(g, s, has, x, ns, ng, mat, r, lo, hi, ulo, uhi, M, runs, staged);
    }
  }
  // shared gather, one owner query at a time: the 64 lanes test the owner's
  // candidate photons, those within the radius are queued in LDS and their
  // terms taken 64 at a time, and a butterfly adds the lanes' partial sums.
  // Every lane of the wave takes part (the caller keeps the whole wave
  // here), so a query's sums do not depend on which other queries share its
  // wave.
  const uint32_t lane = __lane_id();
  const int rank = (int)lane, nAct = 64;
  const unsigned long long ltMask = (1ull << lane) - 1ull;
  V3 myEst = v3(0.0f, 0.0f, 0.0f);
  int myM = 0;
  uint64_t todo = __ballot(has);
  while (todo) {
    const int owner = __ffsll((unsigned long long)todo) - 1;
    todo &= todo - 1ull;
    const V3 c = v3(__shfl(x.x, owner), __shfl(x.y, owner), __shfl(x.z, owner));
    const V3 nso = v3(__shfl(ns.x, owner), __shfl(ns.y, owner), __shfl(ns.z, owner));
    const V3 ngo = v3(__shfl(ng.x, owner), __shfl(ng.y, owner), __shfl(ng.z, owner));
    const float rr = __shfl(r, owner);
    const DevMaterial m = ldgStruct(s.mats + __shfl(mat, owner));
    const float r2 = rr * rr;
    int3 lo = cellOf(c - v3(rr, rr, rr), cs), hi = cellOf(c + v3(rr, rr, rr), cs);
    lo = make_int3(max(lo.x, g.bbox[0]), max(lo.y, g.bbox[1]), max(lo.z, g.bbox[2]));
    hi = make_int3(min(hi.x, g.bbox[3]), min(hi.y, g.bbox[4]), min(hi.z, g.bbox[5]));
    Sum acc;
    int found = 0;
    uint32_t head = 0u, tail = 0u;  // the queue fq[head .. tail) (mod 128) of photons found
    // the cells of the box in x -> y -> z order, 64 at a time: lane rank
    // looks up cell base + rank (all probes in flight together).  Their
    // photons are then dealt to the lanes as one sequence (cell after cell),
    // so every lane has work whatever the cells' sizes.
    const int ny = max(hi.y - lo.y + 1, 0), nz = max(hi.z - lo.z + 1, 0);
    const int nyz = ny * nz, nc = max(hi.x - lo.x + 1, 0) * nyz;
    for (int base = 0; base < nc; base += nAct) {
      const int ci = base + rank;
      uint2 run = make_uint2(0u, 0u);
      if (ci < nc) {
        const int ix = ci / nyz, rem = ci - ix * nyz, iy = rem / nz, iz = rem - iy * nz;
        run = cellLookup(g, packKey(make_int3(lo.x + ix, lo.y + iy, lo.z + iz)));
      }
      // inclusive prefix of the photon counts over the 64 cells
      uint32_t incl = run.y;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (rank >= off) incl += t;
      }
      const uint32_t total = __shfl(incl, 63);
      // sequence index i of this cell's photons maps to photon i + .x
      // (unsigned arithmetic wraps); .y is where the cell ends
      runs[rank] = make_uint2(run.x - (incl - run.y), incl);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      int cur = 0;  // the lane's cell; its sequence indices only grow
      // the trip count is the same on every lane: foundTerms needs the
      // whole wave
      for (uint32_t b0 = 0; b0 < total; b0 += kPhotonLoads * (uint32_t)nAct) {
        uint32_t q[kPhotonLoads];
        float4 xp[kPhotonLoads];
#pragma unroll
        for (uint32_t u = 0; u < kPhotonLoads; ++u) {
          const uint32_t i = b0 + u * (uint32_t)nAct + (uint32_t)rank;
          q[u] = ~0u;
          if (i < total) {
            while (runs[cur].y <= i) ++cur;
            q[u] = runs[cur].x + i;
          }
        }
#pragma unroll
        for (uint32_t u = 0; u < kPhotonLoads; ++u)
          xp[u] = q[u] != ~0u ? g.sX[q[u]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        // photons within the radius join the wave's queue; its terms are
        // taken 64 at a time, one per lane, so the term's arithmetic runs
        // on full waves instead of on the few lanes that found a photon
#pragma unroll
        for (uint32_t u = 0; u < kPhotonLoads; ++u) {
          const V3 dlt = v3(xp[u].x, xp[u].y, xp[u].z) - c;
          const bool in = q[u] != ~0u && dot(dlt, dlt) <= r2;
          const unsigned long long mk = __ballot(in);
          if (in) fq[(tail + (uint32_t)__popcll(mk & ltMask)) & 127u] = q[u];
          tail += (uint32_t)__popcll(mk);
          if (tail - head >= 64u) {
            foundTerms(g, fq, head, 64u, nso, ngo, m, &acc, &found);
            head += 64u;
          }
        }
      }
      __builtin_amdgcn_wave_barrier();  // runs are rewritten for the next cells
    }
    if (tail != head) foundTerms(g, fq, head, tail - head, nso, ngo, m, &acc, &found);
    // the 64 partial sums by a butterfly: each step adds the same two
    // operands on both lanes of a pair (in either order), so every lane ends
    // with the same total
    acc.reduceWave();
    for (int off = 32; off > 0; off >>= 1) found += __shfl_xor(found, off);
    if ((int)lane == owner) {
      myEst = acc.value(g);
      myM = found;
      if (acc.saturated()) atomicAdd(&g.counters[kCtrFluxOverflow], 1u);
    }
  }
  *M = myM;
  return myEst;
}

// Occupancy target of the camera kernel (waves per SIMD; 0: the compiler's
// choice).  At 4 waves it spills ~200 B per lane and is still the fastest
// on the 1920x1080 caustic box (3, 5 and 6 measured slower).
#ifndef HASTY_SPPM_CAMERA_WAVES
#define HASTY_SPPM_CAMERA_WAVES 4
#endif
#if HASTY_SPPM_CAMERA_WAVES > 0
#define HASTY_SPPM_CAMERA_ATTR __attribute__((amdgpu_waves_per_eu(HASTY_SPPM_CAMERA_WAVES)))
#else
#define HASTY_SPPM_CAMERA_ATTR
#endif
// --------------------------------------------------------------- camera ----
// renderPhotonThread's pixel loop (PhotonMapping.cpp:874-925) with
// pathTraceWithCausticsMap (PhotonMapping.cpp:404-503) unrolled into one
// path.  A node's result (photon flux F, light LL, rest LR, specular LS)
// enters the pixel through per-node weights: W_R for F and LR, W_S for LS,
// and a light hit reached through an edge (tD, tC) counts W_S(parent) tC +
// W_R(parent) tD.  Children: W_S' = W_S tC, W_R' = W_R (tD + tC).
__global__ void __launch_bounds__(kB) HASTY_SPPM_CAMERA_ATTR kSppmCamera(DevScene s, SppmDev g, uint32_t pass, float offX, float offY,
                                                  float alpha, int emittedInt, float* color, float* normalOut,
                                                  float* albedoOut) {
  __shared__ uint2 stack[kStackLds * kB];
  uint32_t P = s.width * s.height;
  uint32_t gtid = xcdRemap() * kB + threadIdx.x;
  uint2* ovf = g.ovf + (size_t)gtid * kStackOvf;
  const bool twoStrategies = s.hasLight && s.useMIS;
  for (uint32_t p = gtid; p < P; p += gridDim.x * kB) {
    uint32_t i = p % s.width, j = p / s.width;
    hrng::Stream rng = hrng::make_stream(s.seed, p, pass, hrng::kDomainPath);
    V3 ro, rd;
    cameraRay(s, rng, (float)i + offX, (float)j + offY, &ro, &rd);
    V3 nrm = v3(0.0f, 0.0f, 0.0f), alb = v3(0.0f, 0.0f, 0.0f);
    V3 LL0 = v3(0.0f, 0.0f, 0.0f), LR0 = v3(0.0f, 0.0f, 0.0f), LS0 = v3(0.0f, 0.0f, 0.0f);
    V3 WS = v3(1.0f, 1.0f, 1.0f), WR = v3(1.0f, 1.0f, 1.0f), WLR = v3(0.0f, 0.0f, 0.0f);
    uint32_t nq = 0;
    Media md;
    md.n = 0;
    md.overflow = false;
    md.tx0 = 0u;
    int wavelength = -1;
    Hit h = traceClosest<kStackLds, kStackOvf>(s, ro, rd, stack + threadIdx.x, kB, ovf, nullptr, nullptr);
    for (int depth = 0; depth < s.maxDepth; ++depth) {
      if (h.tri < 0) {
        V3 T = beersLaw(mediaTransmittance(s, md), 1e6f);
        V3 LR = mul(T, evalEnvironment(s, rd));
        if (depth == 0) {
          LR0 = LR;
          nrm = v3(0.0f, 0.0f, 0.0f);
          alb = LR;
        } else {
          LR0 = LR0 + mul(WR, LR);
        }
        break;
      }
      Interaction it = makeInteraction(s, h.tri, h.u, h.v);
      const DevMaterial& m = s.mats[it.mat];
      if (depth == 0) {
        nrm = it.ns;
        alb = albedoOf(m);
      }
      V3 Tb = beersLaw(mediaTransmittance(s, md), norm(it.x - ro));
      V3 wo = -rd;
      if (m.isLight) {
        V3 Le = dot(it.ng, wo) <= 0.0f ? v3(0.0f, 0.0f, 0.0f)
                                        : v3(m.emission[0], m.emission[1], m.emission[2]) * (float)kInvPi;
        V3 LL = mul(Tb, Le);
        if (depth == 0) {
          LL0 = LL;
        } else {
          LS0 = LS0 + mul(WS, LL);
          LR0 = LR0 + mul(WLR, LL);
        }
        break;
      }
      if (m.hasDiffuseLobe) {
        // photon gather of this vertex: recorded for kSppmGather (at most
        // one per depth, so nq < maxDepth)
        const V3 ns = shadingNormal(it, wo);
        float4* qr = g.query + 4 * ((size_t)nq * P + p);
        qr[0] = make_float4(it.x.x, it.x.y, it.x.z, __int_as_float(it.mat));
        qr[1] = make_float4(ns.x, ns.y, ns.z, Tb.x);
        qr[2] = make_float4(it.ng.x, it.ng.y, it.ng.z, Tb.y);
        qr[3] = make_float4(WR.x, WR.y, WR.z, Tb.z);
        ++nq;
      }
      float rrQ = depth >= s.rrStart ? s.rrQ : 0.0f;
      if (rng.uniform01f() <= rrQ) break;
      V3 thr = Tb / (1.0f - rrQ);
      if (depth + 1 >= s.maxDepth) break;  // every child returns zero
      if (twoStrategies) {
        // lightStrategySample (MIS.cpp:32-54, Scene.cpp:538-586)
        float pSurf = !s.hasEnv ? 1.0f : (s.hasSurfaceLight ? 0.5f : 0.0f);
        float xi = rng.uniform01f();
        V3 so, sd;
        int32_t target = -1;
        bool lit = true;  // the sampled light point emits toward so
        if (xi < pSurf) {
          uint32_t li = lightCdfSearch(s, rng.canonical_double());
          target = (int32_t)s.lightTris[li];
          const float* pp = s.triPos + (size_t)target * 9;
          V3 p0 = load3(pp), p1 = load3(pp + 3), p2 = load3(pp + 6);
          float su = rng.uniform01f();
          float tt = hm::fsqrt(rng.uniform01f());
          float b1 = 1.0f - tt;
          float b3 = tt * su;
          float b2 = fmaxf(0.0f, 1.0f - b1 - b3);
          V3 x2 = p0 * b1 + p1 * b2 + p2 * b3;
          so = offsetOrigin(it, dot(it.ng, x2 - it.x) >= 0.0f);
          sd = normalize(x2 - so);
          // the emission test below on the target: seen from behind, zero
          lit = !(dot(normalize(cross(p1 - p0, p2 - p0)), -sd) <= 0.0f);
        } else {
          float dens;
          sd = envSample(s, rng, &dens);
          so = offsetOrigin(it, dot(it.ng, sd) >= 0.0f);
        }
        // a sample the BSDF gives nothing, or that sees a light from behind,
        // adds exactly zero whatever the shadow ray meets: no ray is cast
        const BsdfEval ev = bsdfEvaluate(m, it, wo, sd);
        Hit hl;
        hl.tri = -2;  // not traced
        if (lit && !isZero(ev.fD + ev.fC))
          hl = traceClosest<kStackLds, kStackOvf>(s, so, sd, stack + threadIdx.x, kB, ovf, nullptr, nullptr);
        bool visible = hl.tri != -2 && (target >= 0 ? hl.tri == target
                                                     : (hl.tri < 0 || s.mats[s.triMat[hl.tri]].isLight));
        if (visible) {
          V3 hx = v3(0.0f, 0.0f, 0.0f), hn = v3(0.0f, 0.0f, 0.0f);
          if (hl.tri >= 0) hitPointNormal(s, hl.tri, hl.u, hl.v, &hx, &hn);
          float pL = lightDensity(s, so, sd, hl.tri, hx, hn);
          if (pL != 0.0f) {
            float mm = __builtin_fabsf(dot(sd, it.ng)) / pL;
            V3 sD = ev.fD * mm, sC = ev.fC * mm;
            if (!isZero(sD + sC) && hl.tri != it.tri) {
              float pB = bsdfPdf(m, it, wo, sd, mediaOutsideIOR(s, md, it, wo, wavelength));
              float sum = 0.0f;
              sum += powci<2>(pL);
              sum += powci<2>(pB);
              float w = powci<2>(pL) / sum;
              V3 tD = mul(thr, sD) * w, tC = mul(thr, sC) * w;
              V3 Tc = mediaTransmittanceAfter(s, md, it, wo, sd, 0u, nullptr);
              if (hl.tri >= 0) {
                const DevMaterial& lm = s.mats[s.triMat[hl.tri]];
                V3 Le = dot(hn, -sd) <= 0.0f ? v3(0.0f, 0.0f, 0.0f)
                                              : v3(lm.emission[0], lm.emission[1], lm.emission[2]) * (float)kInvPi;
                V3 LL = mul(beersLaw(Tc, norm(hx - so)), Le);
                LS0 = LS0 + mul(WS, mul(LL, tC));
                LR0 = LR0 + mul(WR, mul(LL, tD));
              } else {
                V3 LR = mul(beersLaw(Tc, 1e6f), evalEnvironment(s, sd));
                LR0 = LR0 + mul(WR, mul(LR, tD + tC));
              }
            }
          }
        }
      }
      // brdfStrategySample (MIS.cpp:6-26, Scene.cpp:328-338)
      float sp = 1.0f;
      int wl = wavelength;
      if (wl == -1 && m.varyIOR) {
        float q0 = 1.0f / 3.0f;
        float xi = rng.uniform01f();
        wl = xi < q0 ? 0 : (xi < q0 + q0 ? 1 : 2);
        sp = q0;
      }
      float outsideIOR = mediaOutsideIOR(s, md, it, wo, wl);
      BsdfSample bs = bsdfSample(m, rng, it, wo, wl, outsideIOR, false, 3);
      bs.tD = bs.tD / sp;
      bs.tC = bs.tC / sp;
      // zero throughput (including pdf 0) ends the path before any trace
      if (bs.pdf == 0.0f || isZero(bs.tD + bs.tC)) break;
      V3 no = offsetOrigin(it, bs.outside);
      Hit h2 = traceClosest<kStackLds, kStackOvf>(s, no, bs.dir, stack + threadIdx.x, kB, ovf, nullptr, nullptr);
      if (h2.tri == it.tri) break;
      float w;
      {
        float pB = bs.pdf;
        float sum = 0.0f;
        if (twoStrategies) {
          V3 hx = v3(0.0f, 0.0f, 0.0f), hn = v3(0.0f, 0.0f, 0.0f);
          if (h2.tri >= 0) hitPointNormal(s, h2.tri, h2.u, h2.v, &hx, &hn);
          float pL = lightDensity(s, no, bs.dir, h2.tri, hx, hn);
          sum += powci<2>(pL);
        }
        sum += powci<2>(pB);
        w = powci<2>(pB) / sum;
      }
      mediaUpdate(md, it, wo, bs.dir, 0u, nullptr);  // photon mapping scenes have no textures
      wavelength = wl;
      V3 tD = mul(thr, bs.tD) * w, tC = mul(thr, bs.tC) * w;
      WLR = mul(WR, tD);
      WS = mul(WS, tC);
      WR = mul(WR, tD + tC);
      ro = no;
      rd = bs.dir;
      h = h2;
    }
    const V3 L = LL0 + LR0 + LS0;
    g.Lpath[p] = L.x, g.Lpath[P + p] = L.y, g.Lpath[2 * P + p] = L.z;
    g.nq[p] = nq;
    normalOut[p * 3] = nrm.x, normalOut[p * 3 + 1] = nrm.y, normalOut[p * 3 + 2] = nrm.z;
    albedoOut[p * 3] = alb.x, albedoOut[p * 3 + 1] = alb.y, albedoOut[p * 3 + 2] = alb.z;
  }
}

// Occupancy target of the gather kernel (waves per SIMD; 0: the compiler's
// choice, 146 VGPRs and 3 waves).  At 4 (128 VGPRs, 40 B of spills) the
// 4M-photon gathers take 67 ms against 69 and the sparse ones 2.7 ms against
// 3.2 on the 1920x1080 caustic box; 5 loses on the dense maps.
#ifndef HASTY_SPPM_GATHER_WAVES
#define HASTY_SPPM_GATHER_WAVES 4
#endif
#if HASTY_SPPM_GATHER_WAVES > 0
#define HASTY_SPPM_GATHER_ATTR __attribute__((amdgpu_waves_per_eu(HASTY_SPPM_GATHER_WAVES)))
#else
#define HASTY_SPPM_GATHER_ATTR
#endif
// The photon gathers of the pass's camera paths (photonMapFluxEstimate,
// PhotonMapping.cpp:271-296, at each diffuse vertex of
// pathTraceWithCausticsMap) and the pixel statistics update.  A wave holds 64
// consecutive pixels; round v gathers every pixel's v-th query, and all 64
// lanes stay in the loop so that the shared gather always has 64 ranks.
// F_R = sum over the vertices in path order of W_R (Tb est), as the camera
// path would have summed it.
template <class Sum>
__global__ void __launch_bounds__(kB) HASTY_SPPM_GATHER_ATTR kSppmGather(DevScene s, SppmDev g, float alpha,
                                                                        int emittedInt, float* color) {
  __shared__ uint2 gatherRuns[kB];
  __shared__ float4 gatherStaged[kB];
  __shared__ uint32_t gatherFound[2 * kB];
  const uint32_t P = s.width * s.height;
  unsigned long long nQueries = 0, nCand = 0, nFound = 0;
  for (uint32_t base = blockIdx.x * kB; base < P; base += gridDim.x * kB) {
    const uint32_t p = base + threadIdx.x;
    const bool valid = p < P;
    float pxRadius = 0.0f;
    uint32_t nq = 0;
    if (valid) {
      pxRadius = g.radius[p];
      if (pxRadius == -1.0f) pxRadius = g.radii[p];
      nq = g.nq[p];
    }
    float radius = pxRadius;
    uint32_t rounds = nq;
    for (int off = 32; off > 0; off >>= 1) rounds = max(rounds, (uint32_t)__shfl_xor((int)rounds, off));
    V3 F0 = v3(0.0f, 0.0f, 0.0f);
    int Mtot = 0;
    for (uint32_t v = 0; v < rounds; ++v) {
      uint32_t cand = 0;
      const bool has = v < nq;
      float4 q0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), q1 = q0, q2 = q0, q3 = q0;
      if (has) {
        const float4* qr = g.query + 4 * ((size_t)v * P + p);
        q0 = qr[0], q1 = qr[1], q2 = qr[2], q3 = qr[3];
      }
      int M = 0;
      const V3 est = gatherQuery<Sum>(g, s, has, f4xyz(q0), f4xyz(q1), f4xyz(q2), __float_as_int(q0.w), radius,
                                      &M, gatherRuns + (threadIdx.x & ~63u), gatherStaged + (threadIdx.x & ~63u),
                                      gatherFound + 2u * (threadIdx.x & ~63u), &cand);
      if (has) {
        nQueries += 1;
        nCand += cand;
        nFound += (unsigned long long)M;
        Mtot += M;
        F0 = F0 + mul(f4xyz(q3), mul(v3(q1.w, q2.w, q3.w), est));
      }
    }
    if (!valid) continue;
    // PMRenderJob::unlockAndUpdateStatistic (PhotonMapping.h:63-105)
    const V3 L = v3(g.Lpath[p], g.Lpath[P + p], g.Lpath[2 * P + p]);
    bool bad = L.x != L.x || L.y != L.y || L.z != L.z || F0.x != F0.x || F0.y != F0.y || F0.z != F0.z;
    V3 LSum = v3(g.LSum[p], g.LSum[P + p], g.LSum[2 * P + p]);
    int Nd = g.Nd[p];
    V3 Lnew;
    if (bad) {
      Lnew = LSum / (1e-7f + (float)Nd);
      g.radius[p] = pxRadius;
    } else {
      LSum = LSum + L;
      Nd += 1;
      float N = g.N[p], Ne = g.Ne[p];
      V3 flux = v3(g.flux[p], g.flux[P + p], g.flux[2 * P + p]);
      V3 Lflux = v3(0.0f, 0.0f, 0.0f);
      if (pxRadius < 0.0f) {
        if (!(radius < 0.0f)) {
          pxRadius = radius;
          N += alpha * Mtot;
          Ne += emittedInt;
          flux = F0;
          Lflux = Ne == 0 ? v3(0.0f, 0.0f, 0.0f) : flux / (pxRadius * pxRadius * (float)kPi * Ne);
        }
      } else {
        float rold = pxRadius;
        float Ni = N;
        if (Ni + Mtot > 0) pxRadius *= hm::fsqrt((Ni + alpha * Mtot) / (Ni + Mtot));
        N += alpha * Mtot;
        Ne += emittedInt;
        flux = (flux + F0) * (pxRadius * pxRadius) / (rold * rold);
        Lflux = Ne == 0 ? v3(0.0f, 0.0f, 0.0f) : flux / (pxRadius * pxRadius * (float)kPi * Ne);
        radius = pxRadius;
      }
      Lnew = Lflux + LSum / (float)Nd;
      g.radius[p] = pxRadius;
      g.N[p] = N;
      g.Ne[p] = Ne;
      g.flux[p] = flux.x, g.flux[P + p] = flux.y, g.flux[2 * P + p] = flux.z;
      g.LSum[p] = LSum.x, g.LSum[P + p] = LSum.y, g.LSum[2 * P + p] = LSum.z;
      g.Nd[p] = Nd;
    }
    g.radii[p] = radius;
    color[p * 3] = Lnew.x, color[p * 3 + 1] = Lnew.y, color[p * 3 + 2] = Lnew.z;
  }
  // one atomic per wave and statistic
  for (int off = 32; off > 0; off >>= 1) {
    nQueries += __shfl_xor(nQueries, off);
    nCand += __shfl_xor(nCand, off);
    nFound += __shfl_xor(nFound, off);
  }
  if ((threadIdx.x & 63u) == 0) {
    atomicAdd(&g.gatherStats[0], nQueries);
    atomicAdd(&g.gatherStats[1], nCand);
    atomicAdd(&g.gatherStats[2], nFound);
  }
}

__global__ void __launch_bounds__(kB) kPixelInit(SppmDev g, uint32_t P) {
  for (uint32_t p = blockIdx.x * kB + threadIdx.x; p < P; p += gridDim.x * kB) {
    g.radius[p] = -1.0f;
    g.N[p] = 0.0f;
    g.Ne[p] = 0.0f;
    g.Nd[p] = 0;
    for (int c = 0; c < 3; ++c) {
      g.flux[c * P + p] = 0.0f;
      g.LSum[c * P + p] = 0.0f;
    }
  }
}
__global__ void __launch_bounds__(kB) kFill(float* a, uint32_t n, float v) {
  for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) a[k] = v;
}
__global__ void kZeroU64(unsigned long long* a, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) a[i] = 0ull;
}
__global__ void kZeroU32(uint32_t* a, uint32_t n) {
  if (threadIdx.x < n) a[threadIdx.x] = 0u;
}

}  // namespace

// ===================================================== SppmRenderer ====
template <class T>
T* SppmRenderer::alloc(size_t n) {
  void* p = nullptr;
  HC(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  allocs_.push_back(p);
  return (T*)p;
}

SppmRenderer::SppmRenderer(const DevScene& scene, const SppmParams& params, hipStream_t stream)
    : s_(scene), p_(params), stream_(stream) {
  W_ = scene.width;
  H_ = scene.height;
  size_t P = (size_t)W_ * H_;
  dPixel_ = alloc<float>(P * 10);
  dRadii_ = alloc<float>(P);
  // gather queries: at most one per path depth and pixel
  dQuery_ = alloc<float4>((size_t)4 * std::max(1, s_.maxDepth) * P);
  dNq_ = alloc<uint32_t>(P);
  dLpath_ = alloc<float>(3 * P);
  dCounters_ = alloc<uint32_t>(8);
  dGatherStats_ = alloc<unsigned long long>(3);
  dCellSize_ = alloc<float>(4);
  dCand_ = alloc<float>(16);
  dState_ = alloc<float>(4);
  dBox_ = alloc<int>(8);
  traceBlocks_ = 2048;
  dOverflow_ = alloc<uint2>((size_t)traceBlocks_ * kB * kStackOvf);
  partialBlocks_ = 512;
  dPartial_ = alloc<float>((size_t)partialBlocks_ * kCandidates);
  // per-thread photon slots: batch x cap x 48 bytes, at most 3 GiB
  cap_ = (uint32_t)std::max(1, s_.maxDepth - 1);
  size_t maxBatch = ((size_t)3 << 30) / ((size_t)cap_ * 48);
  batchMax_ = (uint32_t)std::min<size_t>((size_t)1 << 22, maxBatch);
  batchMax_ = std::max<uint32_t>(batchMax_, 4096);
  dSlX_ = alloc<float4>((size_t)batchMax_ * cap_);
  dSlWi_ = alloc<float4>((size_t)batchMax_ * cap_);
  dSlF_ = alloc<float4>((size_t)batchMax_ * cap_);
  dSlCount_ = alloc<uint32_t>(batchMax_);
  dSlOffset_ = alloc<uint32_t>(batchMax_);
  HC(hipcub::DeviceScan::ExclusiveSum(nullptr, scanTmpBytes_, dSlCount_, dSlOffset_, (int)batchMax_, stream_));
  dScanTmp_ = alloc<char>(scanTmpBytes_);
  // photons: at most storeLimit + cap per pass
  emitLimit_ = (uint32_t)(int)((float)p_.emitFraction * H_ * W_);
  storeLimit_ = (uint32_t)(int)((float)p_.storeFraction * H_ * W_);
  photonCap_ = (size_t)storeLimit_ + cap_ + 1;
  dPhX_ = alloc<float4>(photonCap_);
  dPhWi_ = alloc<float4>(photonCap_);
  dPhF_ = alloc<float4>(photonCap_);
  dSX_ = alloc<float4>(photonCap_);
  dSWi_ = alloc<float4>(photonCap_);
  dSF_ = alloc<float4>(photonCap_);
  dKeys_ = alloc<unsigned long long>(photonCap_);
  dKeysSorted_ = alloc<unsigned long long>(photonCap_);
  dIdx_ = alloc<uint32_t>(photonCap_);
  dIdxSorted_ = alloc<uint32_t>(photonCap_);
  HC(hipcub::DeviceRadixSort::SortPairs(nullptr, sortTmpBytes_, dKeys_, dKeysSorted_, dIdx_, dIdxSorted_,
                                        (int)photonCap_, 0, 3 * kKeyBits, stream_));
  dSortTmp_ = alloc<char>(sortTmpBytes_);
  uint32_t ts = 1024;
  while (ts < 2 * photonCap_) ts <<= 1;
  tableSize_ = ts;
  dTableKeys_ = alloc<unsigned long long>(tableSize_);
  dTableVals_ = alloc<uint2>(tableSize_);
  hipLaunchKernelGGL(kPixelInit, dim3(1024), dim3(kB), 0, stream_,
                     SppmDev{dPixel_, dPixel_ + P, dPixel_ + 2 * P, dPixel_ + 3 * P, dPixel_ + 6 * P,
                             (int*)(dPixel_ + 9 * P), dRadii_},
                     (uint32_t)P);
  HC(hipGetLastError());
}

SppmRenderer::~SppmRenderer() {
  for (void* p : allocs_) (void)hipFree(p);
}

void SppmRenderer::initRadii() {
  size_t P = (size_t)W_ * H_;
  if (!s_.hasSurfaceLight || p_.initialRadius > 0.0f) {
    float r0 = p_.initialRadius > 0.0f ? p_.initialRadius : 1.0f;
    hipLaunchKernelGGL(kFill, dim3(1024), dim3(kB), 0, stream_, dRadii_, (uint32_t)P, r0);
    HC(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(kInitRadii, dim3(traceBlocks_), dim3(kB), 0, stream_, s_, p_.footprintScale, dRadii_,
                     dOverflow_);
  HC(hipGetLastError());
  // getCloseNonNegative (PhotonMapping.cpp:653-678), in raster order and in
  // place as the reference does.
  std::vector<float> r(P);
  HC(hipMemcpyAsync(r.data(), dRadii_, P * sizeof(float), hipMemcpyDeviceToHost, stream_));
  HC(hipStreamSynchronize(stream_));
  int W = (int)W_, H = (int)H_;
  auto inside = [&](int x, int y) { return x >= 0 && y >= 0 && x < W && y < H; };
  auto at = [&](int x, int y) { return r[(size_t)y * W + x]; };
  bool any = false;
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < W; ++i) {
      if (!(r[(size_t)j * W + i] < 0.0f)) continue;
      any = true;
      float v = 1e-3f;
      for (int k = 1; k < std::max(H, W); ++k) {
        bool found = false;
        for (int m = 0; m <= k * 2 && !found; ++m) {
          int i2 = i - k, j2 = j - k + m;
          if (inside(i2, j2) && at(i2, j2) > 0.0f) {
            v = at(i2, j2);
            found = true;
            break;
          }
          i2 = i - k + m;
          j2 = j - k;
          if (inside(i2, j2) && at(i2, j2) > 0.0f) {
            v = at(i2, j2);
            found = true;
            break;
          }
          i2 = i + k;
          j2 = j - k + m;
          if (inside(i2, j2) && at(i2, j2) > 0.0f) {
            v = at(i2, j2);
            found = true;
            break;
          }
        }
        if (found) break;
      }
      r[(size_t)j * W + i] = v;
    }
  if (any) {
    HC(hipMemcpyAsync(dRadii_, r.data(), P * sizeof(float), hipMemcpyHostToDevice, stream_));
    HC(hipStreamSynchronize(stream_));
  }
}

void SppmRenderer::pass(uint32_t passIndex, float offX, float offY, float* dColor, float* dNormal, float* dAlbedo,
                        KernelTimer* timer) {
  size_t P = (size_t)W_ * H_;
  if (!initialised_) {
    if (timer) timer->begin(kKSppmRadii, stream_);
    initRadii();
    if (timer) timer->end(kKSppmRadii, stream_);
    initialised_ = true;
  }
  SppmDev g{};
  g.radius = dPixel_;
  g.N = dPixel_ + P;
  g.Ne = dPixel_ + 2 * P;
  g.flux = dPixel_ + 3 * P;
  g.LSum = dPixel_ + 6 * P;
  g.Nd = (int*)(dPixel_ + 9 * P);
  g.radii = dRadii_;
  g.phX = dPhX_, g.phWi = dPhWi_, g.phF = dPhF_;
  g.sX = dSX_, g.sWi = dSWi_, g.sF = dSF_;
  g.keys = dKeys_;
  g.idx = dIdx_;
  g.tKeys = dTableKeys_;
  g.tVals = dTableVals_;
  g.tMask = tableSize_ - 1;
  g.bbox = dBox_;
  g.cellSize = dCellSize_;
  g.counters = dCounters_;
  g.ovf = dOverflow_;
  g.query = dQuery_;
  g.nq = dNq_;
  g.Lpath = dLpath_;
  g.fluxScale = ldexpf(1.0f, p_.fluxFracBits);
  g.fluxInvScale = ldexp(1.0, -p_.fluxFracBits);
  g.coopCandidates = p_.gatherShareCandidates ? p_.gatherShareCandidates : kCoopCandidatesDefault;
  g.unionFactor = p_.gatherUnionFactor;
  g.gatherStats = dGatherStats_;

  // 1. cell size
  if (timer) timer->begin(kKSppmCellSize, stream_);
  hipLaunchKernelGGL(kCellPartial, dim3(partialBlocks_), dim3(kB), 0, stream_, (const float*)dRadii_, (uint32_t)P,
                     (const float*)dCand_, 0, dPartial_);
  hipLaunchKernelGGL(kCellFold, dim3(1), dim3(kB), 0, stream_, (const float*)dPartial_, (int)partialBlocks_, 0,
                     (uint32_t)P, dCand_, dState_, dCellSize_);
  hipLaunchKernelGGL(kCellPartial, dim3(partialBlocks_), dim3(kB), 0, stream_, (const float*)dRadii_, (uint32_t)P,
                     (const float*)dCand_, 11, dPartial_);
  hipLaunchKernelGGL(kCellFold, dim3(1), dim3(kB), 0, stream_, (const float*)dPartial_, (int)partialBlocks_, 1,
                     (uint32_t)P, dCand_, dState_, dCellSize_);
  hipLaunchKernelGGL(kCellPartial, dim3(partialBlocks_), dim3(kB), 0, stream_, (const float*)dRadii_, (uint32_t)P,
                     (const float*)dCand_, 10, dPartial_);
  hipLaunchKernelGGL(kCellFold, dim3(1), dim3(kB), 0, stream_, (const float*)dPartial_, (int)partialBlocks_, 2,
                     (uint32_t)P, dCand_, dState_, dCellSize_);
  HC(hipGetLastError());
  if (timer) timer->end(kKSppmCellSize, stream_);

  // 2. photons
  hipLaunchKernelGGL(kZeroU32, dim3(1), dim3(64), 0, stream_, dCounters_, 8u);
  hipLaunchKernelGGL(kZeroU64, dim3(1), dim3(64), 0, stream_, dGatherStats_, 3u);
  uint32_t stored = 0, emitted = 0, next = 0;
  if (s_.hasSurfaceLight) {
    uint32_t batch = std::min<uint32_t>(batchMax_, std::max<uint32_t>(65536, storeLimit_));
    while (stored < storeLimit_ && next < emitLimit_) {
      batch = std::min(batch, emitLimit_ - next);
      if (timer) timer->begin(kKSppmPhotons, stream_);
      hipLaunchKernelGGL(kTracePhotons, dim3(traceBlocks_), dim3(kB), 0, stream_, s_, g, next, batch, emitLimit_,
                         passIndex, cap_, dSlX_, dSlWi_, dSlF_, dSlCount_);
      HC(hipGetLastError());
      if (timer) timer->end(kKSppmPhotons, stream_);
      if (timer) timer->begin(kKSppmBuild, stream_);
      size_t tmp = scanTmpBytes_;
      HC(hipcub::DeviceScan::ExclusiveSum(dScanTmp_, tmp, dSlCount_, dSlOffset_, (int)batch, stream_));
      uint32_t blocks = (batch + kB - 1) / kB;
      hipLaunchKernelGGL(kEmitCut, dim3(blocks), dim3(kB), 0, stream_, g, (const uint32_t*)dSlOffset_, next, batch,
                         emitLimit_, storeLimit_);
      hipLaunchKernelGGL(kEmitCompact, dim3(std::min<uint32_t>(blocks, 4096)), dim3(kB), 0, stream_, g,
                         (const uint32_t*)dSlOffset_, (const uint32_t*)dSlCount_, batch, (const float4*)dSlX_,
                         (const float4*)dSlWi_, (const float4*)dSlF_);
      hipLaunchKernelGGL(kEmitAdvance, dim3(1), dim3(1), 0, stream_, g, (const uint32_t*)dSlOffset_,
                         (const uint32_t*)dSlCount_);
      HC(hipGetLastError());
      if (timer) timer->end(kKSppmBuild, stream_);
      uint32_t c[2];
      HC(hipMemcpyAsync(c, dCounters_, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
      HC(hipStreamSynchronize(stream_));
      uint32_t newly = c[1] - emitted;
      stored = c[0];
      emitted = c[1];
      next += batch;
      if (newly < batch) break;  // the emission loop stopped inside this batch
      // size the next batch from the observed store rate
      double rate = emitted > 0 ? (double)stored / (double)emitted : 0.0;
      double want = rate > 0.0 ? 1.25 * (double)(storeLimit_ - std::min(stored, storeLimit_)) / rate : (double)batch * 2;
      batch = (uint32_t)std::min<double>((double)batchMax_, std::max(65536.0, want));
    }
  }
  info_.emitted = emitted;
  info_.stored = stored;

  // 3. cells
  if (timer) timer->begin(kKSppmBuild, stream_);
  hipLaunchKernelGGL(kCellInit, dim3(1024), dim3(kB), 0, stream_, g, 1);
  if (stored > 0) {
    uint32_t blocks = std::min<uint32_t>((stored + kB - 1) / kB, 4096);
    hipLaunchKernelGGL(kCellKeys, dim3(blocks), dim3(kB), 0, stream_, g, stored);
    size_t tmp = sortTmpBytes_;
    HC(hipcub::DeviceRadixSort::SortPairs(dSortTmp_, tmp, dKeys_, dKeysSorted_, dIdx_, dIdxSorted_, (int)stored, 0,
                                          3 * kKeyBits, stream_));
    hipLaunchKernelGGL(kCellBuild, dim3(blocks), dim3(kB), 0, stream_, g, (const unsigned long long*)dKeysSorted_,
                       (const uint32_t*)dIdxSorted_, stored);
    hipLaunchKernelGGL(kCellCount, dim3(blocks), dim3(kB), 0, stream_, g, (const unsigned long long*)dKeysSorted_,
                       stored);
  }
  HC(hipGetLastError());
  if (timer) timer->end(kKSppmBuild, stream_);

  // 4. camera paths (recording their gathers), then the gathers and the
  // statistics
  if (timer) timer->begin(kKSppmCamera, stream_);
  hipLaunchKernelGGL(kSppmCamera, dim3(traceBlocks_), dim3(kB), 0, stream_, s_, g, passIndex, offX, offY, p_.alpha,
                     (int)emitted, dColor, dNormal, dAlbedo);
  HC(hipGetLastError());
  if (timer) timer->end(kKSppmCamera, stream_);
  if (timer) timer->begin(kKSppmGather, stream_);
  const dim3 gatherGrid((uint32_t)std::min<size_t>((P + kB - 1) / kB, 16384));
  if (p_.fixedPointFlux)
    hipLaunchKernelGGL(kSppmGather<FixedSum>, gatherGrid, dim3(kB), 0, stream_, s_, g, p_.alpha, (int)emitted, dColor);
  else
    hipLaunchKernelGGL(kSppmGather<FloatSum>, gatherGrid, dim3(kB), 0, stream_, s_, g, p_.alpha, (int)emitted, dColor);
  HC(hipGetLastError());
  if (timer) timer->end(kKSppmGather, stream_);
  uint32_t c[kNumCtr];
  HC(hipMemcpyAsync(c, dCounters_, sizeof(c), hipMemcpyDeviceToHost, stream_));
  unsigned long long gs[3];
  HC(hipMemcpyAsync(gs, dGatherStats_, sizeof(gs), hipMemcpyDeviceToHost, stream_));
  float cs;
  HC(hipMemcpyAsync(&cs, dCellSize_, sizeof(float), hipMemcpyDeviceToHost, stream_));
  HC(hipStreamSynchronize(stream_));
  info_.cellSize = cs;
  info_.fluxOverflows = c[kCtrFluxOverflow];
  info_.gatherQueries = gs[0];
  info_.gatherCandidates = gs[1];
  info_.gatherFound = gs[2];
  if (c[kCtrKeyOverflow] != 0)
    throw std::runtime_error("photon cell index outside the 21-bit key range (scene extent / cell size too large)");
  passes_ += 1;
}

void SppmRenderer::pixelStats(float* host10) {
  size_t P = (size_t)W_ * H_;
  HC(hipMemcpyAsync(host10, dPixel_, P * 10 * sizeof(float), hipMemcpyDeviceToHost, stream_));
  HC(hipStreamSynchronize(stream_));
}
void SppmRenderer::radii(float* host) {
  size_t P = (size_t)W_ * H_;
  HC(hipMemcpyAsync(host, dRadii_, P * sizeof(float), hipMemcpyDeviceToHost, stream_));
  HC(hipStreamSynchronize(stream_));
}
void SppmRenderer::photons(float* host, size_t maxPhotons) {
  size_t n = std::min<size_t>(info_.stored, maxPhotons);
  std::vector<float4> a(n), b(n), c(n);
  HC(hipMemcpyAsync(a.data(), dPhX_, n * sizeof(float4), hipMemcpyDeviceToHost, stream_));
  HC(hipMemcpyAsync(b.data(), dPhWi_, n * sizeof(float4), hipMemcpyDeviceToHost, stream_));
  HC(hipMemcpyAsync(c.data(), dPhF_, n * sizeof(float4), hipMemcpyDeviceToHost, stream_));
  HC(hipStreamSynchronize(stream_));
  for (size_t k = 0; k < n; ++k) {
    float* o = host + k * 10;
    o[0] = a[k].x, o[1] = a[k].y, o[2] = a[k].z, o[3] = a[k].w;
    o[4] = b[k].x, o[5] = b[k].y, o[6] = b[k].z;
    o[7] = c[k].x, o[8] = c[k].y, o[9] = c[k].z;
  }
}

}  // namespace hasty

// Wavefront path tracer for gfx950: generate -> (extend -> shade -> connect)
// x max_depth -> accumulate, over K samples of every pixel at once.
//
// The recursion of estimateRadiance (core/src/PathTracing.cpp:21-95) becomes
// one extension ray (BSDF strategy) plus one shadow ray (light strategy) per
// vertex: `shade` runs a vertex up to both samples, `connect` finishes the
// light strategy once its closest hit is known (visibility by triangle id,
// Scene.cpp:558-559 / 572-577), and the BSDF strategy's MIS weight is applied
// by the next `shade`, which is where the next hit is known.  Radiance is
// accumulated forward along the path; every decision (roulette, lobe choice,
// rejection, visibility) is made exactly as in the recursion.
//
// Live paths are compacted after every vertex: each wave counts its survivors
// with a ballot, the waves of a block combine their counts with a prefix sum
// in LDS, and the block reserves its range with one atomic.
#include "pt_kernels.h"

namespace hasty {
using namespace hdev;

const char* kernelName(int id) {
  switch (id) {
    case kKGenerate: return "pt_generate";
    case kKExtend: return "pt_extend";
    case kKShade: return "pt_shade";
    case kKConnect: return "pt_connect";
    case kKAccumulate: return "pt_accumulate";
    case kKSppmRadii: return "sppm_init_radii";
    case kKSppmCellSize: return "sppm_cell_size";
    case kKSppmPhotons: return "sppm_photons";
    case kKSppmBuild: return "sppm_build";
    case kKSppmCamera: return "sppm_camera";
    default: return "pt_misc";
  }
}

namespace {

// 8x8 pixel tiles: 64 consecutive slots (one wavefront) cover one tile.
__device__ inline bool slotPixel(uint32_t t, uint32_t width, uint32_t height, uint32_t* i, uint32_t* j) {
  uint32_t tilesX = (width + 7) / 8;
  uint32_t tile = t >> 6, w = t & 63u;
  *i = (tile % tilesX) * 8 + (w & 7u);
  *j = (tile / tilesX) * 8 + (w >> 3);
  return *i < width && *j < height;
}
__device__ inline uint32_t pixelSlot(uint32_t i, uint32_t j, uint32_t width) {
  uint32_t tilesX = (width + 7) / 8;
  return (((j >> 3) * tilesX + (i >> 3)) << 6) | ((j & 7u) << 3) | (i & 7u);
}

// Block-wide append: returns this lane's index in the output queue (valid
// only where `pred`).  Must be reached by every thread of the block.
__device__ inline uint32_t blockAppend(uint32_t* counter, bool pred, uint32_t* ldsWave, uint32_t* ldsBase) {
  uint64_t mask = __ballot(pred);
  uint32_t lane = threadIdx.x & 63u;
  uint32_t wave = threadIdx.x >> 6;
  uint32_t prefix = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0) ldsWave[wave] = (uint32_t)__popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    uint32_t nw = blockDim.x >> 6;
    for (uint32_t w = 0; w < nw; ++w) {
      uint32_t c = ldsWave[w];
      ldsWave[w] = acc;
      acc += c;
    }
    *ldsBase = acc ? atomicAdd(counter, acc) : 0u;
  }
  __syncthreads();
  uint32_t r = *ldsBase + ldsWave[wave] + prefix;
  __syncthreads();
  return r;
}

// Workgroups are dealt round-robin to the 8 XCDs; renumber them so that
// consecutive queue ranges (neighbouring pixels, similar rays) run on the same
// XCD and share its L2.
__device__ inline uint32_t xcdBlock() {
  uint32_t g = gridDim.x, b = blockIdx.x;
  if (g % 8u != 0u) return b;
  return (b % 8u) * (g / 8u) + b / 8u;
}

__device__ inline void waveAddStat(unsigned long long* c, uint32_t v) {
  // lane-sum within the wave, one atomic per wave
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63u) == 0 && v) atomicAdd(c, (unsigned long long)v);
}

__device__ inline void loadMedia(const PathBuffers& pb, uint32_t slot, uint32_t flags, Media& md) {
  uint4 a = pb.media[(size_t)slot * 2], b = pb.media[(size_t)slot * 2 + 1];
  md.e[0] = a.x, md.e[1] = a.y, md.e[2] = a.z, md.e[3] = a.w;
  md.e[4] = b.x, md.e[5] = b.y, md.e[6] = b.z, md.e[7] = b.w;
  md.n = (int)((flags >> 12) & 0xFu);
  md.overflow = (flags >> 16) & 1u;
}
__device__ inline void storeMedia(const PathBuffers& pb, uint32_t slot, const Media& md) {
  pb.media[(size_t)slot * 2] = make_uint4(md.e[0], md.e[1], md.e[2], md.e[3]);
  pb.media[(size_t)slot * 2 + 1] = make_uint4(md.e[4], md.e[5], md.e[6], md.e[7]);
}
__device__ inline uint32_t packFlags(int depth, int wavelength, const Media& md) {
  return (uint32_t)depth | ((uint32_t)(wavelength + 1) << 8) | ((uint32_t)md.n << 12) |
         (md.overflow ? (1u << 16) : 0u);
}

// ------------------------------------------------------------- generate ----
__global__ void __launch_bounds__(kBlock) kGenerate(DevScene s, PathBuffers pb, uint32_t firstSample, uint32_t K,
                                                    const float2* offsets) {
  __shared__ uint32_t ldsWave[kBlock / 64];
  __shared__ uint32_t ldsBase;
  uint32_t Pp = paddedPixels(s.width, s.height);
  uint32_t total = K * Pp;
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
    uint32_t slot = base + threadIdx.x;
    bool live = false;
    if (slot < total) {
      uint32_t k = slot / Pp, t = slot % Pp;
      uint32_t i, j;
      if (slotPixel(t, s.width, s.height, &i, &j)) {
        live = true;
        uint32_t sample = firstSample + k;
        hrng::Stream rng = hrng::make_stream(s.seed, j * s.width + i, sample, hrng::kDomainPath);
        float2 off = offsets[k];
        V3 o, d;
        cameraRay(s, rng, (float)i + off.x, (float)j + off.y, &o, &d);
        pb.rayO[slot] = f4(o, 0.0f);
        pb.rayD[slot] = f4(d, 0.0f);
        pb.beta[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        pb.L[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        pb.auxN[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        pb.auxA[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        pb.state[slot] = make_uint4(rng.counter, 0u, 0xFFFFFFFFu, sample);
      }
    }
    uint32_t idx = blockAppend(&pb.counters[0], live, ldsWave, &ldsBase);
    if (live) pb.qA[idx] = slot;
  }
}

// --------------------------------------------------------------- extend ----
__global__ void __launch_bounds__(kBlock) kExtend(DevScene s, PathBuffers pb, const uint32_t* queue,
                                                  const uint32_t* count, int countStats) {
  __shared__ uint2 stack[kStackLds * kBlock];
  uint32_t n = *count;
  uint32_t gtid = xcdBlock() * blockDim.x + threadIdx.x;
  uint32_t stride = gridDim.x * blockDim.x;
  uint2* ovf = pb.stackOverflow + (size_t)gtid * kStackOvf;
  uint32_t nodes = 0, tris = 0, rays = 0;
  for (uint32_t q = gtid; q < n; q += stride) {
    uint32_t slot = queue[q];
    float4 o = pb.rayO[slot], d = pb.rayD[slot];
    Hit h = traceClosest<kStackLds, kStackOvf>(s, f4xyz(o), f4xyz(d), stack + threadIdx.x, kBlock, ovf,
                                    countStats ? &nodes : nullptr, countStats ? &tris : nullptr);
    pb.hit[slot] = make_float4(h.t, h.u, h.v, __int_as_float(h.tri));
    rays++;
  }
  if (countStats) {
    waveAddStat(&pb.stats[kStatExtRays], rays);
    waveAddStat(&pb.stats[kStatNodes], nodes);
    waveAddStat(&pb.stats[kStatTris], tris);
  }
}

// ---------------------------------------------------------------- shade ----
__global__ void __launch_bounds__(kBlock) kShade(DevScene s, PathBuffers pb, const uint32_t* qIn,
                                                 const uint32_t* countIn, uint32_t* qOut, uint32_t* countOut,
                                                 uint32_t* countShadow) {
  __shared__ uint32_t ldsWave[kBlock / 64];
  __shared__ uint32_t ldsBase;
  uint32_t n = *countIn;
  uint32_t stride = gridDim.x * blockDim.x;
  const bool twoStrategies = s.hasLight && s.useMIS;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    uint32_t q = base + threadIdx.x;
    bool pushNext = false, pushShadow = false;
    uint32_t slot = 0;
    if (q < n) {
      slot = qIn[q];
      uint4 st = pb.state[slot];
      int depth = (int)(st.y & 0xFFu);
      int wavelength = (int)((st.y >> 8) & 0x3u) - 1;
      Media md;
      loadMedia(pb, slot, st.y, md);
      int32_t prevTri = (int32_t)st.z;
      float4 ro4 = pb.rayO[slot], rd4 = pb.rayD[slot], h4 = pb.hit[slot];
      V3 ro = f4xyz(ro4), rd = f4xyz(rd4);
      int32_t tri = __float_as_int(h4.w);
      float4 b4 = pb.beta[slot];
      V3 beta = f4xyz(b4);
      float4 L4 = pb.L[slot];
      V3 L = f4xyz(L4);
      hrng::Stream rng;
      {
        uint32_t Pp = paddedPixels(s.width, s.height);
        uint32_t i, j;
        slotPixel(slot % Pp, s.width, s.height, &i, &j);
        rng = hrng::make_stream(s.seed, j * s.width + i, st.w, hrng::kDomainPath);
        rng.counter = st.x;
      }
      bool alive = true;
      // The BSDF sample that produced this ray: self-intersection, then MIS.
      if (depth > 0) {
        if (tri == prevTri) {
          alive = false;
        } else {
          float pB = rd4.w;
          float w;
          if (twoStrategies) {
            V3 hx = v3(0.0f, 0.0f, 0.0f), hn = v3(0.0f, 0.0f, 0.0f);
            if (tri >= 0) hitPointNormal(s, tri, h4.y, h4.z, &hx, &hn);
            float pL = lightDensity(s, ro, rd, tri, hx, hn);
            float sum = 0.0f;
            sum += powci<2>(pL);
            sum += powci<2>(pB);
            w = powci<2>(pB) / sum;
          } else {
            float sum = 0.0f;
            sum += powci<2>(pB);
            w = powci<2>(pB) / sum;
          }
          beta = beta * w;
        }
      }
      if (alive) {
        float4 aux = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (tri < 0) {
          V3 T = beersLaw(mediaTransmittance(s, md), 1e6f);
          V3 Lc = mul(T, evalEnvironment(s, rd));
          L = L + mul(beta, Lc);
          if (depth == 0) {
            pb.auxN[slot] = aux;
            pb.auxA[slot] = f4(Lc, 0.0f);
          }
          alive = false;
        } else {
          Interaction it = makeInteraction(s, tri, h4.y, h4.z);
          const DevMaterial& m = s.mats[it.mat];
          if (depth == 0) {
            pb.auxN[slot] = f4(it.ns, 0.0f);
            pb.auxA[slot] = make_float4(m.albedo[0], m.albedo[1], m.albedo[2], 0.0f);
          }
          V3 Tb = beersLaw(mediaTransmittance(s, md), norm(it.x - ro));
          V3 wo = -rd;
          if (m.isLight) {
            V3 Le = dot(it.ng, wo) <= 0.0f ? v3(0.0f, 0.0f, 0.0f)
                                            : v3(m.emission[0], m.emission[1], m.emission[2]) * (float)kInvPi;
            L = L + mul(beta, mul(Tb, Le));
            alive = false;
          } else {
            float rrQ = depth >= s.rrStart ? s.rrQ : 0.0f;
            if (rng.uniform01f() < rrQ) {
              alive = false;
            } else {
              V3 thr = Tb / (1.0f - rrQ);
              if (depth + 1 >= s.maxDepth) {
                alive = false;  // every deeper estimate is zero
              } else {
                if (twoStrategies) {
                  // ---- light strategy (MIS.cpp:32-54, Scene.cpp:538-586)
                  float pSurf = !s.hasEnv ? 1.0f : (s.hasSurfaceLight ? 0.5f : 0.0f);
                  float xi = rng.uniform01f();
                  V3 so, sd;
                  int32_t target = -1;
                  if (xi < pSurf) {
                    uint32_t li = lightCdfSearch(s, rng.canonical_double());
                    target = (int32_t)s.lightTris[li];
                    const float* p = s.triPos + (size_t)target * 9;
                    V3 p0 = load3(p), p1 = load3(p + 3), p2 = load3(p + 6);
                    float su = rng.uniform01f();
                    float tt = hm::fsqrt(rng.uniform01f());
                    float b1 = 1.0f - tt;
                    float b3 = tt * su;
                    float b2 = fmaxf(0.0f, 1.0f - b1 - b3);
                    V3 x2 = p0 * b1 + p1 * b2 + p2 * b3;
                    V3 diff = x2 - it.x;
                    so = offsetOrigin(it, dot(it.ng, diff) >= 0.0f);
                    sd = normalize(x2 - so);
                  } else {
                    float dens;
                    sd = envSample(s, rng, &dens);
                    so = offsetOrigin(it, dot(it.ng, sd) >= 0.0f);
                  }
                  BsdfEval ev = bsdfEvaluate(m, it, wo, sd);
                  float outsideIOR = mediaOutsideIOR(s, md, it, wo, wavelength);
                  float pB = bsdfPdf(m, it, wo, sd, outsideIOR);
                  Media mc = md;
                  mediaUpdate(mc, it, wo, sd);
                  V3 Tc = mediaTransmittance(s, mc);
                  pb.shO[slot] = f4(so, pB);
                  pb.shD[slot] = f4(sd, __int_as_float(target));
                  pb.shFD[slot] = f4(ev.fD, __builtin_fabsf(dot(sd, it.ng)));
                  pb.shFC[slot] = f4(ev.fC, __int_as_float(tri));
                  pb.shThr[slot] = f4(thr, 0.0f);
                  pb.shBeta[slot] = f4(beta, 0.0f);
                  pb.shT[slot] = f4(Tc, 0.0f);
                  pushShadow = true;
                }
                // ---- BSDF strategy (MIS.cpp:6-26, Scene.cpp:328-338)
                float sampleProbability = 1.0f;
                if (wavelength == -1 && m.varyIOR) {
                  float p0 = 1.0f / 3.0f;
                  float xi = rng.uniform01f();
                  if (xi < p0) wavelength = 0;
                  else if (xi < p0 + p0) wavelength = 1;
                  else wavelength = 2;
                  sampleProbability = p0;
                }
                float outsideIOR = mediaOutsideIOR(s, md, it, wo, wavelength);
                BsdfSample bs = bsdfSample(m, rng, it, wo, wavelength, outsideIOR, false, 3);
                bs.tD = bs.tD / sampleProbability;
                bs.tC = bs.tC / sampleProbability;
                if (bs.pdf != 0.0f) {
                  V3 t2 = mul(thr, bs.tD + bs.tC);
                  if (!isZero(t2)) {
                    V3 no = offsetOrigin(it, bs.outside);
                    mediaUpdate(md, it, wo, bs.dir);
                    pb.rayO[slot] = f4(no, 0.0f);
                    pb.rayD[slot] = f4(bs.dir, bs.pdf);
                    pb.beta[slot] = f4(mul(beta, t2), 0.0f);
                    pushNext = true;
                  }
                }
                alive = false;
                if (pushNext) {
                  pb.state[slot] = make_uint4(rng.counter, packFlags(depth + 1, wavelength, md), (uint32_t)tri, st.w);
                  storeMedia(pb, slot, md);
                }
              }
            }
          }
        }
      }
      pb.L[slot] = f4(L, 0.0f);
    }
    uint32_t idxN = blockAppend(countOut, pushNext, ldsWave, &ldsBase);
    if (pushNext) qOut[idxN] = slot;
    uint32_t idxS = blockAppend(countShadow, pushShadow, ldsWave, &ldsBase);
    if (pushShadow) pb.qS[idxS] = slot;
  }
}

// -------------------------------------------------------------- connect ----
__global__ void __launch_bounds__(kBlock) kConnect(DevScene s, PathBuffers pb, const uint32_t* count, int countStats) {
  __shared__ uint2 stack[kStackLds * kBlock];
  uint32_t n = *count;
  uint32_t gtid = xcdBlock() * blockDim.x + threadIdx.x;
  uint32_t stride = gridDim.x * blockDim.x;
  uint2* ovf = pb.stackOverflow + (size_t)gtid * kStackOvf;
  uint32_t nodes = 0, tris = 0, rays = 0;
  for (uint32_t q = gtid; q < n; q += stride) {
    uint32_t slot = pb.qS[q];
    float4 o4 = pb.shO[slot], d4 = pb.shD[slot];
    V3 so = f4xyz(o4), sd = f4xyz(d4);
    int32_t target = __float_as_int(d4.w);
    RayPre rp = prepareRay(so, sd);
    rays++;
    // Visibility exactly as the closest-hit test of Scene.cpp:558-559 / 572-577,
    // answered with an early-exit search wherever that is equivalent.
    Hit h;
    bool visible;
    if (target >= 0) {
      // closest hit == target  <=>  the target is hit at (tT, target) and no
      // other triangle hits before it (or at tT with a smaller index)
      const float* p = s.triPos + (size_t)target * 9;
      float4 p0 = make_float4(p[0], p[1], p[2], 0.0f), p1 = make_float4(p[3], p[4], p[5], 0.0f),
             p2 = make_float4(p[6], p[7], p[8], 0.0f);
      float tT, uT, vT;
      visible = intersectTri(rp, p0, p1, p2, 1e-6f, &tT, &uT, &vT);
      if (visible) {
        visible = findBlocker<kStackLds, kStackOvf>(s, rp, tT, target, stack + threadIdx.x, kBlock, ovf,
                                                    countStats ? &nodes : nullptr, countStats ? &tris : nullptr) < 0;
        h.t = tT, h.u = uT, h.v = vT, h.tri = target;
      }
    } else if (!s.hasSurfaceLight) {
      // no emissive triangle: visible <=> nothing is hit at all
      visible = findBlocker<kStackLds, kStackOvf>(s, rp, __builtin_inff(), 0x7FFFFFFF, stack + threadIdx.x, kBlock,
                                                  ovf, countStats ? &nodes : nullptr,
                                                  countStats ? &tris : nullptr) < 0;
      h.t = __builtin_inff(), h.u = h.v = 0.0f, h.tri = -1;
    } else {
      h = traceClosest<kStackLds, kStackOvf>(s, so, sd, stack + threadIdx.x, kBlock, ovf,
                                             countStats ? &nodes : nullptr, countStats ? &tris : nullptr);
      bool hitLight = h.tri >= 0 && s.mats[s.triMat[h.tri]].isLight;
      visible = h.tri < 0 || hitLight;
    }
    if (!visible) continue;
    V3 hx = v3(0.0f, 0.0f, 0.0f), hn = v3(0.0f, 0.0f, 0.0f);
    if (h.tri >= 0) hitPointNormal(s, h.tri, h.u, h.v, &hx, &hn);
    float pL = lightDensity(s, so, sd, h.tri, hx, hn);
    if (pL == 0.0f) continue;
    float4 fd4 = pb.shFD[slot], fc4 = pb.shFC[slot];
    float mm = fd4.w / pL;
    V3 tD = f4xyz(fd4) * mm, tC = f4xyz(fc4) * mm;
    V3 thr = f4xyz(pb.shThr[slot]);
    V3 t2 = mul(thr, tD + tC);
    if (isZero(t2) || h.tri == __float_as_int(fc4.w)) continue;
    float pB = o4.w;
    float sum = 0.0f;
    sum += powci<2>(pL);
    sum += powci<2>(pB);
    float w = powci<2>(pL) / sum;
    t2 = t2 * w;
    V3 Tc = f4xyz(pb.shT[slot]);
    V3 Li;
    if (h.tri < 0) {
      Li = mul(beersLaw(Tc, 1e6f), evalEnvironment(s, sd));
    } else {
      const DevMaterial& lm = s.mats[s.triMat[h.tri]];
      V3 Tb = beersLaw(Tc, norm(hx - so));
      V3 Le = dot(hn, -sd) <= 0.0f ? v3(0.0f, 0.0f, 0.0f)
                                    : v3(lm.emission[0], lm.emission[1], lm.emission[2]) * (float)kInvPi;
      Li = mul(Tb, Le);
    }
    V3 beta = f4xyz(pb.shBeta[slot]);
    float4 L4 = pb.L[slot];
    pb.L[slot] = f4(f4xyz(L4) + mul(beta, mul(Li, t2)), 0.0f);
  }
  if (countStats) {
    waveAddStat(&pb.stats[kStatShadowRays], rays);
    waveAddStat(&pb.stats[kStatNodes], nodes);
    waveAddStat(&pb.stats[kStatTris], tris);
  }
}

// ----------------------------------------------------------- accumulate ----
__global__ void __launch_bounds__(kBlock) kAccumulate(DevScene s, PathBuffers pb, uint32_t K, float* color,
                                                      float* normal, float* albedo) {
  uint32_t P = s.width * s.height;
  uint32_t Pp = paddedPixels(s.width, s.height);
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += stride) {
    uint32_t i = p % s.width, j = p / s.width;
    uint32_t t = pixelSlot(i, j, s.width);
    float c0 = color[p * 3], c1 = color[p * 3 + 1], c2 = color[p * 3 + 2];
    float n0 = normal[p * 3], n1 = normal[p * 3 + 1], n2 = normal[p * 3 + 2];
    float a0 = albedo[p * 3], a1 = albedo[p * 3 + 1], a2 = albedo[p * 3 + 2];
    for (uint32_t k = 0; k < K; ++k) {
      uint32_t slot = k * Pp + t;
      float4 L = pb.L[slot];
      if (L.x != L.x || L.y != L.y || L.z != L.z) continue;  // PathTracing.cpp:132-135
      float4 an = pb.auxN[slot], aa = pb.auxA[slot];
      c0 += L.x, c1 += L.y, c2 += L.z;
      n0 += an.x, n1 += an.y, n2 += an.z;
      a0 += aa.x, a1 += aa.y, a2 += aa.z;
    }
    color[p * 3] = c0, color[p * 3 + 1] = c1, color[p * 3 + 2] = c2;
    normal[p * 3] = n0, normal[p * 3 + 1] = n1, normal[p * 3 + 2] = n2;
    albedo[p * 3] = a0, albedo[p * 3 + 1] = a1, albedo[p * 3 + 2] = a2;
  }
}

__global__ void kResetCounters(uint32_t* c) {
  if (threadIdx.x < kNumCounters) c[threadIdx.x] = 0u;
}
// Next bounce: qB becomes the input queue.
__global__ void kAdvance(uint32_t* c) {
  if (threadIdx.x == 0) {
    c[0] = c[1];
    c[1] = 0u;
    c[2] = 0u;
  }
}
// --------------------------------------------------------- trace only ----
__global__ void __launch_bounds__(kBlock) kTraceRays(DevScene s, const float* rays, uint32_t n, int32_t* tri,
                                                     float* tuv, uint2* overflow) {
  __shared__ uint2 stack[kStackLds * kBlock];
  uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t stride = gridDim.x * blockDim.x;
  uint2* ovf = overflow + (size_t)gtid * kStackOvf;
  for (uint32_t q = gtid; q < n; q += stride) {
    const float* r = rays + (size_t)q * 6;
    Hit h = traceClosest<kStackLds, kStackOvf>(s, v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), stack + threadIdx.x, kBlock, ovf,
                                    nullptr, nullptr);
    tri[q] = h.tri;
    tuv[(size_t)q * 3] = h.t;
    tuv[(size_t)q * 3 + 1] = h.u;
    tuv[(size_t)q * 3 + 2] = h.v;
  }
}

}  // namespace

void launchPTBatch(const DevScene& s, const PathBuffers& pbIn, uint32_t firstSample, uint32_t K, const float2* dOffsets,
                   float* dColor, float* dNormal, float* dAlbedo, const LaunchConfig& cfg, hipStream_t stream,
                   KernelTimer* timer) {
  PathBuffers pb = pbIn;  // queue pointers swap per bounce
  uint32_t Pp = paddedPixels(s.width, s.height);
  uint32_t total = K * Pp;
  auto t0 = [&](int k) {
    if (timer) timer->begin(k, stream);
  };
  auto t1 = [&](int k) {
    if (timer) timer->end(k, stream);
  };
  uint32_t genBlocks = (total + kBlock - 1) / kBlock;
  if (genBlocks > cfg.shadeBlocks) genBlocks = cfg.shadeBlocks;
  t0(kKMisc);
  kResetCounters<<<1, 64, 0, stream>>>(pb.counters);
  t1(kKMisc);
  t0(kKGenerate);
  kGenerate<<<genBlocks, kBlock, 0, stream>>>(s, pb, firstSample, K, dOffsets);
  t1(kKGenerate);
  int countStats = cfg.countStats ? 1 : 0;
  for (int d = 0; d < s.maxDepth; ++d) {
    t0(kKExtend);
    kExtend<<<cfg.traceBlocks, kBlock, 0, stream>>>(s, pb, pb.qA, &pb.counters[0], countStats);
    t1(kKExtend);
    t0(kKShade);
    kShade<<<genBlocks, kBlock, 0, stream>>>(s, pb, pb.qA, &pb.counters[0], pb.qB, &pb.counters[1], &pb.counters[2]);
    t1(kKShade);
    t0(kKConnect);
    kConnect<<<cfg.traceBlocks, kBlock, 0, stream>>>(s, pb, &pb.counters[2], countStats);
    t1(kKConnect);
    t0(kKMisc);
    kAdvance<<<1, 64, 0, stream>>>(pb.counters);
    t1(kKMisc);
    // swap queues: the next extension reads what shade wrote
    uint32_t* tmp = pb.qA;
    pb.qA = pb.qB;
    pb.qB = tmp;
  }
  uint32_t P = s.width * s.height;
  uint32_t accBlocks = (P + kBlock - 1) / kBlock;
  t0(kKAccumulate);
  kAccumulate<<<accBlocks, kBlock, 0, stream>>>(s, pb, K, dColor, dNormal, dAlbedo);
  t1(kKAccumulate);
}

void launchTraceClosest(const DevScene& s, const float* dRays, uint32_t n, int32_t* dTri, float* dTuv, uint2* dOverflow,
                        uint32_t overflowThreads, hipStream_t stream) {
  uint32_t blocks = (n + kBlock - 1) / kBlock;
  uint32_t maxBlocks = overflowThreads / kBlock;
  if (blocks > maxBlocks) blocks = maxBlocks;
  if (blocks == 0) return;
  kTraceRays<<<blocks, kBlock, 0, stream>>>(s, dRays, n, dTri, dTuv, dOverflow);
}

}  // namespace hasty

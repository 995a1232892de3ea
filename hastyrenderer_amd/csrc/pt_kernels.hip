// Wavefront path tracer for gfx950 with path regeneration:
//   fill -> repeat { extend -> shade -> connect -> retire } -> accumulate.
// A render call covers n samples of every pixel (n * pixels paths, numbered
// sample-major, pixels in 8x8-tile order).  A fixed pool of path slots runs
// them: `retire` writes a finished path's radiance into a results ring and
// starts the next path id in the same slot, so the queues stay full until
// the last paths drain.  Every path's random stream is keyed by (pixel,
// sample), so the image does not depend on which slot ran which path, and
// `accumulate` adds each pixel's samples in sample order.
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

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

namespace hasty {
using namespace hdev;

const char* kernelName(int id) {
  switch (id) {
    case kKGenerate: return "pt_generate";
    case kKExtend: return "pt_extend";
    case kKShade: return "pt_shade";
    case kKConnect: return "pt_connect";
    case kKAccumulate: return "pt_accumulate";
    case kKRetire: return "pt_retire";
    case kKGraph: return "pt_graph_chunk";
    case kKGuideInfer: return "guide_infer";
    case kKSppmRadii: return "sppm_init_radii";
    case kKSppmCellSize: return "sppm_cell_size";
    case kKSppmPhotons: return "sppm_photons";
    case kKSppmBuild: return "sppm_build";
    case kKSppmCamera: return "sppm_camera";
    case kKSppmGather: return "sppm_gather";
    default: return "pt_misc";
  }
}

namespace {

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

// Entries beyond md.n are never read, so an empty stack is neither loaded
// nor stored.
__device__ inline void loadMedia(const PathBuffers& pb, uint32_t slot, uint32_t flags, const uint32_t* txs,
                                 Media& md) {
  md.n = (int)((flags >> 12) & 0xFu);
  md.overflow = (flags >> 16) & 1u;
  md.tx0 = 0u;
  if (md.n == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) md.e[i] = 0u;
    return;
  }
  uint4 a = pb.media[(size_t)slot * 2], b = pb.media[(size_t)slot * 2 + 1];
  md.e[0] = a.x, md.e[1] = a.y, md.e[2] = a.z, md.e[3] = a.w;
  md.e[4] = b.x, md.e[5] = b.y, md.e[6] = b.z, md.e[7] = b.w;
  if (txs) md.tx0 = txs[0];
}
__device__ inline void storeMedia(const PathBuffers& pb, uint32_t slot, const Media& md) {
  pb.media[(size_t)slot * 2] = make_uint4(md.e[0], md.e[1], md.e[2], md.e[3]);
  pb.media[(size_t)slot * 2 + 1] = make_uint4(md.e[4], md.e[5], md.e[6], md.e[7]);
}
__device__ inline uint32_t packFlags(int depth, int wavelength, const Media& md) {
  return (uint32_t)depth | ((uint32_t)(wavelength + 1) << 8) | ((uint32_t)md.n << 12) |
         (md.overflow ? (1u << 16) : 0u);
}

// ------------------------------------------------------------------ jobs ----
// Job of a global path id: the last job whose pidBase <= pid.
template <bool MULTI>
__device__ inline uint32_t findJob(const JobSet& js, uint32_t pid) {
  if (!MULTI) return 0u;
  uint32_t lo = 0, hi = js.nJobs;
  while (hi - lo > 1u) {
    uint32_t mid = (lo + hi) >> 1;
    if (js.passes[mid].pidBase <= pid) lo = mid;
    else hi = mid;
  }
  return lo;
}
// Batched rendering keeps the job index of a path in flag bits 23-31 (set
// when the path starts) for up to 512 jobs; larger batches search.
constexpr uint32_t kJobShift = 23;
constexpr uint32_t kFlagJobs = 512;
template <bool MULTI>
__device__ inline uint32_t jobOf(const JobSet& js, const uint4& st) {
  if (!MULTI) return 0u;
  return js.nJobs <= kFlagJobs ? st.y >> kJobShift : findJob<true>(js, st.w);
}
__device__ inline uint32_t jobBits(const JobSet& js, uint32_t job) {
  return js.nJobs <= kFlagJobs ? job << kJobShift : 0u;
}
// A copy of *p, which every active lane of the wave passes, read through
// the scalar unit.
template <class T>
__device__ inline T loadUniform(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "word-sized");
  typedef __attribute__((address_space(4))) const uint32_t CWord;
  const CWord* w = (const CWord*)p;
  uint32_t words[sizeof(T) / 4];
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) words[i] = w[i];
  T out;
  __builtin_memcpy(&out, words, sizeof(T));
  return out;
}
// A pointer read from memory is generic to the compiler (flat loads, which
// wait on the LDS counter too); these round trips through the global address
// space give every scene and pass array global loads.
template <class T>
__device__ inline const T* globalPtr(const T* p) {
  return (const T*)(const __attribute__((address_space(1))) T*)p;
}
__device__ inline void makeGlobal(DevScene& s) {
  s.nodes = globalPtr(s.nodes), s.tris = globalPtr(s.tris), s.triPos = globalPtr(s.triPos);
  s.triNrm = globalPtr(s.triNrm), s.triMat = globalPtr(s.triMat), s.triGeom = globalPtr(s.triGeom);
  s.triSmooth = globalPtr(s.triSmooth), s.mats = globalPtr(s.mats), s.lightTris = globalPtr(s.lightTris);
  s.lightCdf = globalPtr(s.lightCdf), s.envRGB = globalPtr(s.envRGB), s.envPdf = globalPtr(s.envPdf);
  s.envCdf = globalPtr(s.envCdf), s.envCdfGuide = globalPtr(s.envCdfGuide);
}
__device__ inline void makeGlobal(PassInfo& pi) {
  pi.tileOrder = globalPtr(pi.tileOrder), pi.rankOf = globalPtr(pi.rankOf), pi.offsets = globalPtr(pi.offsets);
}
// f(scene, pass) of the lane's job.  Batched, the wave's active lanes are
// served one distinct job at a time (normally one: paths are handed out in
// id order), each with its scene and pass in scalar registers.
template <bool MULTI, class F>
__device__ inline void forJob(const JobSet& js, uint32_t job, F&& f) {
  if constexpr (!MULTI) {
    f(js.s, js.pi);
  } else {
    for (;;) {
      const uint32_t j0 = __builtin_amdgcn_readfirstlane(job);
      if (job == j0) {
        DevScene sc = loadUniform(js.scenes + j0);
        PassInfo pi = loadUniform(js.passes + j0);
        makeGlobal(sc);
        makeGlobal(pi);
        f(sc, pi);
        break;
      }
    }
  }
}
// Pixel and sample of a path (global id) of job pi.
__device__ inline void pathPixel(const PassInfo& pi, uint32_t pid, uint32_t* pixel, uint32_t* k) {
  uint32_t local = pid - pi.pidBase;
  uint32_t kk = local / pi.pixels;
  *pixel = pi.tileOrder[local - kk * pi.pixels];
  *k = kk;
}

// ------------------------------------------------------------ path start ----
// Camera ray of path `pid` (Camera.cpp:41-81 with the sample's offset,
// RenderJob.h:53-60) into `slot`.
__device__ inline void startPath(const DevScene& s, const PathBuffers& pb, const PassInfo& pi, uint32_t slot,
                                 uint32_t pid, uint32_t flags) {
  uint32_t k, pixel;
  pathPixel(pi, pid, &pixel, &k);
  uint32_t i = pixel % s.width, j = pixel / s.width;
  hrng::Stream rng = hrng::make_stream(s.seed, pixel, pi.first + k, hrng::kDomainPath);
  float2 off = pi.offsets[k];
  V3 o, d;
  cameraRay(s, rng, (float)i + off.x, (float)j + off.y, &o, &d);
  // throughput, radiance and the depth-0 normal / albedo are not stored
  // here: kShade starts every path from throughput 1 and radiance 0 and
  // writes the aux values at its first vertex
  pb.rayO[slot] = f4(o, 0.0f);
  pb.rayD[slot] = f4(d, 0.0f);
  pb.state[slot] = make_uint4(rng.counter, flags, 0xFFFFFFFFu, pid);
}

// Every slot idle, no path started.
// Occupancy targets (waves per SIMD) of the three ray-bound stages: the
// register budget each may use is 512 / waves VGPRs.  0 leaves it to the
// compiler.  Extend at 8 waves (64 VGPRs) is ~7% faster than at the
// compiler's 7; shade at 3 keeps the guided variant (which would take 190
// VGPRs) at the plain one's occupancy without spills (4 and 5 lose); connect
// at 5 (96 VGPRs, 28 B of spills) is ~1% faster than the compiler's 4, at 6
// slower.
#ifndef HASTY_EXTEND_WAVES
#define HASTY_EXTEND_WAVES 8
#endif
#ifndef HASTY_SHADE_WAVES
#define HASTY_SHADE_WAVES 3
#endif
#ifndef HASTY_CONNECT_WAVES
#define HASTY_CONNECT_WAVES 5
#endif
#define HASTY_WAVES_ATTR_(n) __attribute__((amdgpu_waves_per_eu(n)))
#if HASTY_EXTEND_WAVES > 0
#define HASTY_EXTEND_ATTR HASTY_WAVES_ATTR_(HASTY_EXTEND_WAVES)
#else
#define HASTY_EXTEND_ATTR
#endif
#if HASTY_SHADE_WAVES > 0
#define HASTY_SHADE_ATTR HASTY_WAVES_ATTR_(HASTY_SHADE_WAVES)
#else
#define HASTY_SHADE_ATTR
#endif
#if HASTY_CONNECT_WAVES > 0
#define HASTY_CONNECT_ATTR HASTY_WAVES_ATTR_(HASTY_CONNECT_WAVES)
#else
#define HASTY_CONNECT_ATTR
#endif

// The paths to extend and shade this bounce: while they fill more than half
// the pool they are visited in slot order (entry q is slot q, taken if its
// mark is this bounce), so that neighbouring lanes read neighbouring path
// records; otherwise through the queue qA.  Either way every path is visited
// once, and the image does not depend on the order.
struct ExtSet {
  bool dense;
  uint32_t total, epoch;
  __device__ bool member(const PathBuffers& pb, uint32_t q, uint32_t* slot) const {
    if (!dense) {
      *slot = pb.qA[q];
      return true;
    }
    *slot = q;
    return pb.mark[q] == epoch;
  }
};
__device__ inline ExtSet extensionSet(const PathBuffers& pb) {
  ExtSet e;
  const uint32_t n = pb.counters[kCntA];
  e.dense = 2ull * n > (unsigned long long)pb.capacity;
  e.total = e.dense ? pb.capacity : n;
  e.epoch = pb.counters[kCntEpoch];
  return e;
}

__global__ void __launch_bounds__(kBlock) kInitRender(PathBuffers pb, uint32_t limit) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < pb.capacity; k += stride) pb.mark[k] = 0xFFFFFFFFu;
  if (blockIdx.x == 0 && threadIdx.x < kNumCounters) {
    uint32_t v = 0u;
    if (threadIdx.x == kCntLimit) v = limit;
    pb.counters[threadIdx.x] = v;
  }
}

// A slot is idle when its mark is not this bounce's (its path ended, or it
// never had one).  kFill visits the slots in order: idle slots draw offsets
// from the fill cursor (one atomic per block), those below the ring-space
// limit start the paths next + offset and join qA; kFillCommit advances
// next by the number started and resets the cursor.  Paths are started in
// slot order, so their records are written in whole cache lines where idle
// slots are adjacent.
__device__ inline uint32_t fillAvail(const uint32_t* c) {
  const uint32_t next = c[kCntNext], lim = c[kCntLimit];
  return lim > next ? lim - next : 0u;
}
// Block-wide sum of v (every thread of the block calls it).
__device__ inline uint32_t blockSum(uint32_t v, uint32_t* ldsWave) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63u) == 0) ldsWave[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t sum = 0;
  for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) sum += ldsWave[w];
  __syncthreads();
  return sum;
}
// Exclusive prefix count of pred in thread order over the block, and the
// block's count in *total (every thread of the block calls it).
__device__ inline uint32_t blockPrefix(bool pred, uint32_t* ldsWave, uint32_t* total) {
  const uint64_t mask = __ballot(pred);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t prefix = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0) ldsWave[wave] = (uint32_t)__popcll(mask);
  __syncthreads();
  uint32_t before = 0, sum = 0;
  for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) {
    const uint32_t c = ldsWave[w];
    before += w < wave ? c : 0u;
    sum += c;
  }
  __syncthreads();
  *total = sum;
  return before + prefix;
}
// Each block owns a contiguous range of slots: it counts the range's idle
// slots, takes offsets for all of them with one atomic on the fill cursor
// and a qA range for those it can start with one more, then starts them in
// slot order.  (Two atomics per block, instead of two per 256 slots: same-
// address atomics from every block were what bounded this kernel.)  8192
// blocks (1024 slots each at 2^23 slots) measured fastest: 2048 and 4096
// leave too few waves to hide the mark reads, 16384 and more pay for the
// atomics again.
constexpr uint32_t kFillBlocks = 8192;
template <bool MULTI>
__global__ void __launch_bounds__(kBlock) kFill(JobSet js, PathBuffers pb) {
  __shared__ uint32_t ldsWave[kBlock / 64];
  __shared__ uint32_t ldsBase[3];
  __shared__ uint32_t ldsIdle[2 * kBlock];  // idle slots waiting for a path, in slot order
  const uint32_t avail = fillAvail(pb.counters);
  if (avail == 0u) return;  // the same for every thread
  const uint32_t next = pb.counters[kCntNext], epoch = pb.counters[kCntEpoch];
  const uint32_t per = (pb.capacity + gridDim.x * kBlock - 1) / (gridDim.x * kBlock) * kBlock;
  const uint32_t lo = min(blockIdx.x * per, pb.capacity), hi = min(lo + per, pb.capacity);
  uint32_t mine = 0;
  for (uint32_t slot = lo + threadIdx.x; slot < hi; slot += kBlock) mine += pb.mark[slot] != epoch ? 1u : 0u;
  const uint32_t idleTotal = blockSum(mine, ldsWave);
  if (threadIdx.x == 0) {
    const uint32_t b = idleTotal ? atomicAdd(&pb.counters[kCntFill], idleTotal) : 0u;
    const uint32_t live = b < avail ? min(idleTotal, avail - b) : 0u;
    ldsBase[0] = b;
    ldsBase[1] = live;
    ldsBase[2] = live ? atomicAdd(&pb.counters[kCntA], live) : 0u;
  }
  __syncthreads();
  const uint32_t b = ldsBase[0], live = ldsBase[1], qa = ldsBase[2];
  // The r-th idle slot of the range starts path next + b + r.  The idle
  // slots are queued in LDS and their paths started a block at a time, one
  // per thread, so that the camera-ray code runs on full waves.
  auto start = [&](uint32_t r, uint32_t slot) {
    const uint32_t pid = next + b + r;
    const uint32_t j = findJob<MULTI>(js, pid);
    const uint32_t flags = MULTI ? jobBits(js, j) : 0u;
    pb.mark[slot] = epoch;
    forJob<MULTI>(js, j, [&](const DevScene& sc, const PassInfo& pi) { startPath(sc, pb, pi, slot, pid, flags); });
    pb.qA[qa + r] = slot;
  };
  uint32_t rankBase = 0, done = 0;  // idle slots seen / started (uniform)
  for (uint32_t base = lo; base < hi && rankBase < live; base += kBlock) {  // uniform bounds
    const uint32_t slot = base + threadIdx.x;
    const bool idle = slot < hi && pb.mark[slot] != epoch;
    uint32_t count;
    const uint32_t r = rankBase + blockPrefix(idle, ldsWave, &count);
    if (idle && r < live) ldsIdle[r - done] = slot;  // fewer than kBlock were queued before
    rankBase += count;
    const uint32_t queued = min(rankBase, live) - done;
    if (queued >= kBlock) {
      __syncthreads();
      start(done + threadIdx.x, ldsIdle[threadIdx.x]);
      const uint32_t rest = threadIdx.x < queued - kBlock ? ldsIdle[kBlock + threadIdx.x] : 0u;
      __syncthreads();
      if (threadIdx.x < queued - kBlock) ldsIdle[threadIdx.x] = rest;
      done += kBlock;
    }
  }
  __syncthreads();
  const uint32_t queued = min(rankBase, live) - done;
  if (threadIdx.x < queued) start(done + threadIdx.x, ldsIdle[threadIdx.x]);
}
__global__ void kFillCommit(uint32_t* c) {
  if (threadIdx.x == 0) {
    const uint32_t avail = fillAvail(c), idle = c[kCntFill];
    c[kCntNext] += idle < avail ? idle : avail;
    c[kCntFill] = 0u;
  }
}
__global__ void kSetLimit(uint32_t* c, uint32_t limit) {
  if (threadIdx.x == 0) c[kCntLimit] = limit;
}

// --------------------------------------------------------------- extend ----
// COUNT: add rays, node visits and triangle tests to the statistics (the
// counting is compiled out otherwise).
// W: node width of the single job's scene (4 or 8), 0 = each job's own.
template <bool MULTI, bool COUNT, int W>
__global__ void __launch_bounds__(kBlock) HASTY_EXTEND_ATTR kExtend(JobSet js, PathBuffers pb) {
  __shared__ uint2 stack[kStackLds * kBlock];
  const ExtSet es = extensionSet(pb);
  uint32_t gtid = xcdBlock() * blockDim.x + threadIdx.x;
  uint32_t stride = gridDim.x * blockDim.x;
  uint2* ovf = pb.stackOverflow + (size_t)gtid * kStackOvf;
  uint32_t nodes = 0, tris = 0, rays = 0;
  for (uint32_t q = gtid; q < es.total; q += stride) {
    uint32_t slot;
    if (!es.member(pb, q, &slot)) continue;
    float4 o = pb.rayO[slot], d = pb.rayD[slot];
    const uint32_t job = MULTI ? jobOf<MULTI>(js, pb.state[slot]) : 0u;
    Hit h;
    forJob<MULTI>(js, job, [&](const DevScene& sc, const PassInfo&) {
      h = traceClosest<kStackLds, kStackOvf, W>(sc, f4xyz(o), f4xyz(d), stack + threadIdx.x, kBlock, ovf, &nodes,
                                                &tris);
    });
    pb.hit[slot] = make_float4(h.t, h.u, h.v, __int_as_float(h.tri));
    rays++;
  }
  if (COUNT) {
    waveAddStat(&pb.stats[kStatExtRays], rays);
    waveAddStat(&pb.stats[kStatNodes], nodes);
    waveAddStat(&pb.stats[kStatTris], tris);
  }
}

// ---------------------------------------------------------------- shade ----
// Bits 20-21 of the flags: training records of the path (guided rendering).
constexpr uint32_t kRecShift = 20;
// Materials guided by the learned proposal: GGX alpha at least this.
constexpr float kGuideMinAlpha = 0.25f;
// Vertices that sample the guide: rough, non-emitting reflectors.  A
// roughness texture makes the material eligible whatever its texels (the
// one-sample mixture stays unbiased at glossy texels).
__device__ inline bool guideEligible(const DevMaterial& m) {
  return !m.isLight && m.transmission == 0.0f && (m.tex[1] >= 0 || m.alpha >= kGuideMinAlpha);
}
// A path whose throughput has grown past g.maxThroughput samples its BSDF
// alone from then on.  Which technique samples a vertex may depend on the
// path so far without bias; this keeps the guide's 1 / (1 - fraction)
// weights from compounding along long chains of guided vertices.
__device__ inline bool guideThroughputOk(const GuideRun& g, V3 beta) {
  return fmaxf(beta.x, fmaxf(beta.y, beta.z)) <= g.maxThroughput;
}

template <bool MULTI, bool GUIDED>
__global__ void __launch_bounds__(kBlock) HASTY_SHADE_ATTR kShade(JobSet js, PathBuffers pb) {
  __shared__ uint32_t ldsWave[kBlock / 64];
  __shared__ uint32_t ldsBase;
  const ExtSet es = extensionSet(pb);
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < es.total; base += stride) {
    uint32_t q = base + threadIdx.x;
    bool pushNext = false, pushShadow = false, live = false;
    uint32_t slot = 0;
    if (q < es.total && es.member(pb, q, &slot)) {
      live = true;
      const uint4 st = pb.state[slot];
      forJob<MULTI>(js, jobOf<MULTI>(js, st), [&](const DevScene& s, const PassInfo& pi) {
      const bool twoStrategies = s.hasLight && s.useMIS;
      int depth = (int)(st.y & 0xFFu);
      int wavelength = (int)((st.y >> 8) & 0x3u) - 1;
      // per-entry albedo texels of the media stack (albedo-textured scenes)
      uint32_t* txs = s.texturedMedia ? pb.mediaTx + (size_t)slot * kMaxMedia : nullptr;
      Media md;
      loadMedia(pb, slot, st.y, txs, md);
      int32_t prevTri = (int32_t)st.z;
      float4 ro4 = pb.rayO[slot], rd4 = pb.rayD[slot], h4 = pb.hit[slot];
      V3 ro = f4xyz(ro4), rd = f4xyz(rd4);
      int32_t tri = __float_as_int(h4.w);
      // a path's first vertex starts from throughput 1 and radiance 0 (not
      // stored by kFill)
      const float4 b4 = depth == 0 ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : pb.beta[slot];
      V3 beta = f4xyz(b4);
      const float4 L4 = depth == 0 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : pb.L[slot];
      V3 L = f4xyz(L4);
      uint32_t nrec = GUIDED ? (st.y >> kRecShift) & 3u : 0u;
      if (GUIDED && depth > 0 && nrec > 0 && ((st.y >> (kRecShift + 2)) & 1u)) {
        // radiance before the continuation of the last recorded vertex
        js.g.rec[(size_t)slot * 12 + 6 * (nrec - 1) + 5] = L4;
      }
      hrng::Stream rng;
      {
        uint32_t k, pixel;
        pathPixel(pi, st.w, &pixel, &k);
        rng = hrng::make_stream(s.seed, pixel, pi.first + k, hrng::kDomainPath);
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
          DevMaterial m = ldgStruct(s.mats + it.mat);
          // guided vertex: rough reflection (no delta lobe, and not so glossy
          // that broad vMF lobes could only waste samples); decided on the
          // material before its textures and on the stored throughput
          // before this vertex's MIS weight, exactly as kGuideInfer decides
          // it, so that every guided vertex has its mixture table
          const bool guided = GUIDED && guideEligible(m) && (uint32_t)depth < js.g.maxDepth &&
                              (depth == 0 || guideThroughputOk(js.g, f4xyz(b4)));
          if (depth == 0) pb.auxN[slot] = f4(it.ns, 0.0f);  // the default shading normal
          uint32_t texel = 0u;
          if (m.textured) applyTextures(s, m, it, 1.0f - h4.y - h4.z, h4.y, h4.z, &texel);
          if (depth == 0) pb.auxA[slot] = make_float4(m.albedo[0], m.albedo[1], m.albedo[2], 0.0f);
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
                // the vertex's mixture table, written by kGuideInfer
                const float4* gm = GUIDED ? js.g.raw + (size_t)slot * 10 : nullptr;
                // kGuideInfer wrote the learned probability of guided vertices
                const float gf = GUIDED ? (js.g.sel && guided ? js.g.sel[slot] : js.g.fraction) : 0.0f;
                if (twoStrategies) {
                  // ---- light strategy (MIS.cpp:32-54, Scene.cpp:538-586)
                  float pSurf = !s.hasEnv ? 1.0f : (s.hasSurfaceLight ? 0.5f : 0.0f);
                  float xi = rng.uniform01f();
                  V3 so, sd;
                  int32_t target = -1;
                  bool lit = true;  // the sampled light point emits toward so
                  if (xi < pSurf) {
                    uint32_t li = lightCdfSearch(s, rng.canonical_double());
                    target = (int32_t)ldg(s.lightTris + li);
                    const float* p = s.triPos + (size_t)target * 9;
                    V3 p0 = load3g(p), p1 = load3g(p + 3), p2 = load3g(p + 6);
                    float su = rng.uniform01f();
                    float tt = hm::fsqrt(rng.uniform01f());
                    float b1 = 1.0f - tt;
                    float b3 = tt * su;
                    float b2 = fmaxf(0.0f, 1.0f - b1 - b3);
                    V3 x2 = p0 * b1 + p1 * b2 + p2 * b3;
                    V3 diff = x2 - it.x;
                    so = offsetOrigin(it, dot(it.ng, diff) >= 0.0f);
                    sd = normalize(x2 - so);
                    // kConnect's emission test on the target (hitPointNormal's
                    // normal): a one-sided light seen from behind gives zero
                    lit = !(dot(normalize(cross(p1 - p0, p2 - p0)), -sd) <= 0.0f);
                  } else {
                    float dens;
                    sd = envSample(s, rng, &dens);
                    so = offsetOrigin(it, dot(it.ng, sd) >= 0.0f);
                  }
                  BsdfEval ev = bsdfEvaluate(m, it, wo, sd);
                  // A light sample the BSDF gives nothing (e.g. below the
                  // surface; kConnect's t2 = thr (fD + fC) |cos| / pL) or
                  // that sees a light from behind adds exactly zero whatever
                  // its shadow ray meets, so no ray is cast; about a quarter
                  // of the samples on the GGX grid.
                  if (lit && !isZero(ev.fD + ev.fC)) {
                    float outsideIOR = mediaOutsideIOR(s, md, it, wo, wavelength);
                    float pB = bsdfPdf(m, it, wo, sd, outsideIOR);
                    if (GUIDED && guided) pB = gf * guide::mixturePdfTable(gm, sd) + (1.0f - gf) * pB;
                    V3 Tc = mediaTransmittanceAfter(s, md, it, wo, sd, texel, txs);
                    pb.shO[slot] = f4(so, pB);
                    pb.shD[slot] = f4(sd, __int_as_float(target));
                    pb.shFD[slot] = f4(ev.fD, __builtin_fabsf(dot(sd, it.ng)));
                    pb.shFC[slot] = f4(ev.fC, __int_as_float(tri));
                    pb.shThr[slot] = f4(thr, 0.0f);
                    pb.shBeta[slot] = f4(beta, 0.0f);
                    pb.shT[slot] = f4(Tc, 0.0f);
                    pushShadow = true;
                  }
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
                BsdfSample bs;
                float pBsdf = 0.0f;  // the BSDF's density of the direction (training records)
                if (GUIDED && guided) {
                  // one-sample MIS of guide and BSDF: p = f pG + (1 - f) pB
                  if (rng.uniform01f() < gf) {
                    float x0 = rng.uniform01f(), x1 = rng.uniform01f(), x2 = rng.uniform01f();
                    bs.dir = guide::mixtureSampleTable(gm, x0, x1, x2);
                    bs.outside = dot(it.ng, wo) >= 0.0f;
                    BsdfEval ev = bsdfEvaluate(m, it, wo, bs.dir);
                    float pB = bsdfPdf(m, it, wo, bs.dir, outsideIOR);
                    pBsdf = pB;
                    bs.pdf = gf * guide::mixturePdfTable(gm, bs.dir) + (1.0f - gf) * pB;
                    float mm = bs.pdf > 0.0f ? __builtin_fabsf(dot(bs.dir, it.ng)) / bs.pdf : 0.0f;
                    bs.tD = ev.fD * mm;
                    bs.tC = ev.fC * mm;
                    if (!(bs.pdf > 0.0f)) bs.pdf = 0.0f;
                  } else {
                    bs = bsdfSample(m, rng, it, wo, wavelength, outsideIOR, false, 3);
                    pBsdf = bs.pdf;
                    if (bs.pdf != 0.0f) {
                      float pMix = gf * guide::mixturePdfTable(gm, bs.dir) + (1.0f - gf) * bs.pdf;
                      float sc = bs.pdf / pMix;
                      bs.tD = bs.tD * sc;
                      bs.tC = bs.tC * sc;
                      bs.pdf = pMix;
                    }
                  }
                } else {
                  bs = bsdfSample(m, rng, it, wo, wavelength, outsideIOR, false, 3);
                }
                bs.tD = bs.tD / sampleProbability;
                bs.tC = bs.tC / sampleProbability;
                if (bs.pdf != 0.0f) {
                  V3 t2 = mul(thr, bs.tD + bs.tC);
                  if (!isZero(t2)) {
                    V3 no = offsetOrigin(it, bs.outside);
                    mediaUpdate(md, it, wo, bs.dir, texel, txs);
                    pb.rayO[slot] = f4(no, 0.0f);
                    pb.rayD[slot] = f4(bs.dir, bs.pdf);
                    pb.beta[slot] = f4(mul(beta, t2), 0.0f);
                    pushNext = true;
                  }
                }
                alive = false;
                if (pushNext) {
                  uint32_t recBits = 0u;
                  if (GUIDED) {
                    if (guided && nrec < 2u && js.g.sampleCap > 0u) {  // records only when training
                      float4* rr = js.g.rec + (size_t)slot * 12 + 6 * nrec;
                      V3 nf = dot(it.ng, wo) >= 0.0f ? it.ng : -it.ng;
                      rr[0] = f4(it.x, 0.0f);
                      rr[1] = f4(nf, 0.0f);
                      rr[2] = f4(wo, bs.pdf);
                      rr[3] = f4(bs.dir, pBsdf);
                      rr[4] = f4(beta, 0.0f);
                      nrec += 1u;
                      recBits = (nrec << kRecShift) | (1u << (kRecShift + 2));
                    } else {
                      recBits = nrec << kRecShift;
                    }
                  }
                  pb.mark[slot] = es.epoch + 1u;
                  pb.state[slot] =
                      make_uint4(rng.counter,
                                 packFlags(depth + 1, wavelength, md) | recBits | (st.y & ~((1u << kJobShift) - 1u)),
                                 (uint32_t)tri, st.w);
                  if (md.n > 0) storeMedia(pb, slot, md);
                }
              }
            }
          }
        }
      }
      pb.L[slot] = f4(L, 0.0f);
      });
    }
    uint32_t idxN = blockAppend(&pb.counters[kCntB], pushNext, ldsWave, &ldsBase);
    if (pushNext) pb.qB[idxN] = slot;
    uint32_t idxS = blockAppend(&pb.counters[kCntS], pushShadow, ldsWave, &ldsBase);
    if (pushShadow) pb.qS[idxS] = slot;
    bool retire = live && !pushNext;  // the path ends here (after its shadow ray, if any)
    uint32_t idxR = blockAppend(&pb.counters[kCntR], retire, ldsWave, &ldsBase);
    if (retire) pb.qR[idxR] = slot;
  }
}

// -------------------------------------------------------------- connect ----
template <bool MULTI, bool COUNT, int W>
__global__ void __launch_bounds__(kBlock) HASTY_CONNECT_ATTR kConnect(JobSet js, PathBuffers pb) {
  __shared__ uint2 stack[kStackLds * kBlock];
  uint32_t n = pb.counters[kCntS];
  uint32_t gtid = xcdBlock() * blockDim.x + threadIdx.x;
  uint32_t stride = gridDim.x * blockDim.x;
  uint2* ovf = pb.stackOverflow + (size_t)gtid * kStackOvf;
  uint32_t nodes = 0, tris = 0, rays = 0;
  for (uint32_t q = gtid; q < n; q += stride) {
    uint32_t slot = pb.qS[q];
    const uint32_t job = MULTI ? jobOf<MULTI>(js, pb.state[slot]) : 0u;
    forJob<MULTI>(js, job, [&](const DevScene& s, const PassInfo&) {
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
      float4 p0 = make_float4(ldg(p), ldg(p + 1), ldg(p + 2), 0.0f),
             p1 = make_float4(ldg(p + 3), ldg(p + 4), ldg(p + 5), 0.0f),
             p2 = make_float4(ldg(p + 6), ldg(p + 7), ldg(p + 8), 0.0f);
      float tT, uT, vT;
      visible = intersectTri(rp, p0, p1, p2, 1e-6f, &tT, &uT, &vT);
      if (visible) {
        visible = findBlocker<kStackLds, kStackOvf, W>(s, rp, tT, target, stack + threadIdx.x, kBlock, ovf,
                                                    &nodes, &tris) < 0;
        h.t = tT, h.u = uT, h.v = vT, h.tri = target;
      }
    } else if (!s.hasSurfaceLight) {
      // no emissive triangle: visible <=> nothing is hit at all
      visible = findBlocker<kStackLds, kStackOvf, W>(s, rp, __builtin_inff(), 0x7FFFFFFF, stack + threadIdx.x, kBlock,
                                                  ovf, &nodes,
                                                  &tris) < 0;
      h.t = __builtin_inff(), h.u = h.v = 0.0f, h.tri = -1;
    } else {
      h = traceClosest<kStackLds, kStackOvf, W>(s, so, sd, stack + threadIdx.x, kBlock, ovf,
                                             &nodes, &tris);
      bool hitLight = h.tri >= 0 && ldg(&s.mats[ldg(s.triMat + h.tri)].isLight);
      visible = h.tri < 0 || hitLight;
    }
    if (!visible) return;
    V3 hx = v3(0.0f, 0.0f, 0.0f), hn = v3(0.0f, 0.0f, 0.0f);
    if (h.tri >= 0) hitPointNormal(s, h.tri, h.u, h.v, &hx, &hn);
    float pL = lightDensity(s, so, sd, h.tri, hx, hn);
    if (pL == 0.0f) return;
    float4 fd4 = pb.shFD[slot], fc4 = pb.shFC[slot];
    float mm = fd4.w / pL;
    V3 tD = f4xyz(fd4) * mm, tC = f4xyz(fc4) * mm;
    V3 thr = f4xyz(pb.shThr[slot]);
    V3 t2 = mul(thr, tD + tC);
    if (isZero(t2) || h.tri == __float_as_int(fc4.w)) return;
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
      const DevMaterial lm = ldgStruct(s.mats + ldg(s.triMat + h.tri));
      V3 Tb = beersLaw(Tc, norm(hx - so));
      V3 Le = dot(hn, -sd) <= 0.0f ? v3(0.0f, 0.0f, 0.0f)
                                    : v3(lm.emission[0], lm.emission[1], lm.emission[2]) * (float)kInvPi;
      Li = mul(Tb, Le);
    }
    V3 beta = f4xyz(pb.shBeta[slot]);
    float4 L4 = pb.L[slot];
    pb.L[slot] = f4(f4xyz(L4) + mul(beta, mul(Li, t2)), 0.0f);
    });
  }
  if (COUNT) {
    waveAddStat(&pb.stats[kStatShadowRays], rays);
    waveAddStat(&pb.stats[kStatShadowNodes], nodes);
    waveAddStat(&pb.stats[kStatShadowTris], tris);
  }
}

// --------------------------------------------------------------- retire ----
// Finished paths: results into the ring; the slot is idle from the next
// bounce on (its mark is not renewed) and kFill gives it a new path id.
// With `count`, paths whose media stack overflowed are counted.
template <bool GUIDED>
__global__ void __launch_bounds__(kBlock) kRetire(JobSet js, PathBuffers pb, uint32_t count) {
  uint32_t n = pb.counters[kCntR];
  uint32_t stride = gridDim.x * blockDim.x;
  uint32_t overflowed = 0;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const uint32_t slot = pb.qR[q];
    uint4 st = pb.state[slot];
    overflowed += (st.y >> 16) & 1u;
    uint32_t ri = st.w % pb.ringPaths;
    float4 Lf = pb.L[slot];
    pb.ringL[ri] = Lf;
    if (GUIDED) {
      // training samples: weight = luminance of the radiance the
      // continuation brought, over the throughput before the sample
      uint32_t nrec = (st.y >> kRecShift) & 3u;
      for (uint32_t r = 0; r < nrec; ++r) {
        const float4* rr = js.g.rec + (size_t)slot * 12 + 6 * r;
        float4 b = rr[4], ls = rr[5];
        float cx = b.x > 0.0f ? (Lf.x - ls.x) / b.x : 0.0f;
        float cy = b.y > 0.0f ? (Lf.y - ls.y) / b.y : 0.0f;
        float cz = b.z > 0.0f ? (Lf.z - ls.z) / b.z : 0.0f;
        float F = 0.2126f * cx + 0.7152f * cy + 0.0722f * cz;
        if (F > 0.0f && F < 3.0e38f) {
          uint32_t k = atomicAdd(js.g.sampleCount, 1u);
          if (k < js.g.sampleCap) {
            float4 x = rr[0];
            js.g.samples[4 * (size_t)k] = make_float4(x.x, x.y, x.z, F);
            js.g.samples[4 * (size_t)k + 1] = rr[1];
            js.g.samples[4 * (size_t)k + 2] = rr[2];
            js.g.samples[4 * (size_t)k + 3] = rr[3];
          }
        }
      }
    }
    pb.ringN[ri] = pb.auxN[slot];
    pb.ringA[ri] = pb.auxA[slot];
  }
  if (count) waveAddStat(&pb.stats[kStatMediaOverflow], overflowed);
}

// ----------------------------------------------------------- accumulate ----
// Samples [k0, k1) of job pi, from the ring, added pixel by pixel in sample
// order; NaN samples are skipped (PathTracing.cpp:132-135).
__global__ void __launch_bounds__(kBlock) kAccumulate(PathBuffers pb, PassInfo pi, uint32_t k0, uint32_t k1,
                                                      float* color, float* normal, float* albedo) {
  uint32_t P = pi.pixels;
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += stride) {
    float c0 = color[p * 3], c1 = color[p * 3 + 1], c2 = color[p * 3 + 2];
    float n0 = normal[p * 3], n1 = normal[p * 3 + 1], n2 = normal[p * 3 + 2];
    float a0 = albedo[p * 3], a1 = albedo[p * 3 + 1], a2 = albedo[p * 3 + 2];
    uint32_t pid0 = pi.pidBase + pi.rankOf[p];
    for (uint32_t k = k0; k < k1; ++k) {
      uint32_t ri = (pid0 + k * P) % pb.ringPaths;
      float4 L = pb.ringL[ri];
      if (L.x != L.x || L.y != L.y || L.z != L.z) continue;
      float4 an = pb.ringN[ri], aa = pb.ringA[ri];
      c0 += L.x, c1 += L.y, c2 += L.z;
      n0 += an.x, n1 += an.y, n2 += an.z;
      a0 += aa.x, a1 += aa.y, a2 += aa.z;
    }
    color[p * 3] = c0, color[p * 3 + 1] = c1, color[p * 3 + 2] = c2;
    normal[p * 3] = n0, normal[p * 3 + 1] = n1, normal[p * 3 + 2] = n2;
    albedo[p * 3] = a0, albedo[p * 3 + 1] = a1, albedo[p * 3 + 2] = a2;
  }
}

// ------------------------------------------------------------ guide infer ----
// Raw guide outputs for every path in qA that hit a guided surface: inputs
// are the hit point, the geometric normal facing wo, and wo.  A wave takes 64
// paths, one per lane, for the gathers (hit, material, triangle, ray), then
// runs the network on its four 16-path tiles in turn; the tile's lanes fetch
// their path's inputs from the gathering lane by shuffles.  Blocks of 512
// threads, three per CU (the LDS copy of the network is 41 KiB per block).
#ifndef HASTY_GUIDE_INFER_WAVES
#define HASTY_GUIDE_INFER_WAVES 6
#endif
constexpr uint32_t kGuideInferThreads = 512, kGuideInferBlocks = 256 * 3;
__global__ void __launch_bounds__(kGuideInferThreads) HASTY_WAVES_ATTR_(HASTY_GUIDE_INFER_WAVES)
    kGuideInfer(JobSet js, PathBuffers pb) {
  __shared__ guide::LdsNet ldsNet;
  guide::stageNet(js.g.net, &ldsNet, true);
  const DevScene& s = js.s;
  const uint32_t n = pb.counters[kCntA];
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t g = wave; g * 64 < n; g += waves) {
    const uint32_t e = g * 64 + lane;
    uint32_t slot = 0u;
    bool want = false;
    V3 x = v3(0.0f, 0.0f, 0.0f), ng = v3(0.0f, 0.0f, 1.0f), wo = v3(0.0f, 0.0f, 1.0f);
    if (e < n) {
      slot = pb.qA[e];
      const float4 h4 = pb.hit[slot];
      const int32_t tri = __float_as_int(h4.w);
      if (tri >= 0) {
        // only hits on guided materials read the outputs
        const DevMaterial m = ldgStruct(s.mats + ldg(s.triMat + tri));
        const uint32_t depth = pb.state[slot].y & 0xFFu;
        want = guideEligible(m) && depth < js.g.maxDepth;
        if (want && depth > 0) want = guideThroughputOk(js.g, f4xyz(pb.beta[slot]));
        if (want) {
          hitPointNormal(s, tri, h4.y, h4.z, &x, &ng);
          wo = -f4xyz(pb.rayD[slot]);
          if (dot(ng, wo) < 0.0f) ng = -ng;
          x = guide::normPos(js.g.net, x);
        }
      }
    }
    const unsigned long long wantMask = __ballot(want);
    for (int j = 0; j < 4; ++j) {
      // a tile whose 16 paths have no guided hit skips the network (the test
      // is uniform over the wave)
      if (((wantMask >> (16 * j)) & 0xFFFFull) == 0ull) continue;
      const int src = 16 * j + r;
      const V3 xq = v3(__shfl(x.x, src), __shfl(x.y, src), __shfl(x.z, src));
      const V3 nq = v3(__shfl(ng.x, src), __shfl(ng.y, src), __shfl(ng.z, src));
      const V3 wq = v3(__shfl(wo.x, src), __shfl(wo.y, src), __shfl(wo.z, src));
      const uint32_t sq = __shfl(slot, src);
      // lane group q holds lobes 2q and 2q + 1 of the path's mixture
      float r10[10], w0, w1, logit;
      guide::tileOutputs(&ldsNet, xq, nq, wq, r10, &w0, &w1, &logit);
      if ((wantMask >> src) & 1ull) {
        float4* out = js.g.raw + (size_t)sq * 10;
        out[2 * q] = guide::lobeEntry(r10, w0);
        out[2 * q + 1] = guide::lobeEntry(r10 + 5, w1);
        reinterpret_cast<float2*>(out + guide::kLobes)[q] = make_float2(w0, w1);
        if (js.g.sel && q == 0) js.g.sel[sq] = guide::selectionOf(js.g.net, logit);
      }
    }
  }
}

// Next bounce: qB becomes the input queue (the host swaps the pointers).
__global__ void kAdvance(uint32_t* c) {
  if (threadIdx.x == 0) {
    c[kCntEpoch] += 1u;
    c[kCntA] = c[kCntB];
    c[kCntB] = 0u;
    c[kCntS] = 0u;
    c[kCntR] = 0u;
  }
}

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

// Closest-hit records (HitRecord) of n rays, grid-stride over the rays.
__global__ void __launch_bounds__(kBlock) kTraceHits(DevScene s, const float* rays, uint32_t n, HitRecord* hits,
                                                     uint2* overflow) {
  __shared__ uint2 stack[kStackLds * kBlock];
  uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t stride = gridDim.x * blockDim.x;
  uint2* ovf = overflow + (size_t)gtid * kStackOvf;
  for (uint32_t q = gtid; q < n; q += stride) {
    const float* r = rays + (size_t)q * 6;
    Hit h = traceClosest<kStackLds, kStackOvf>(s, v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), stack + threadIdx.x,
                                                kBlock, ovf, nullptr, nullptr);
    HitRecord rec;
    if (h.tri < 0) {
      rec.geomID = rec.primID = 0xFFFFFFFFu;
      rec.t = __builtin_inff();
      rec.u = rec.v = 0.0f;
      rec.ng[0] = rec.ng[1] = rec.ng[2] = 0.0f;
    } else {
      const uint32_t g = ldg(s.triGeom + h.tri);
      rec.geomID = g;
      rec.primID = (uint32_t)h.tri - ldg(s.shapeTriOffset + g);
      rec.t = h.t;
      rec.u = h.u;
      rec.v = h.v;
      const float* p = s.triPos + (size_t)h.tri * 9;
      const V3 p0 = load3g(p), p1 = load3g(p + 3), p2 = load3g(p + 6);
      const V3 ng = normalize(cross(p1 - p0, p2 - p0));  // Scene.cpp:297, Embree's Ng orientation
      rec.ng[0] = ng.x;
      rec.ng[1] = ng.y;
      rec.ng[2] = ng.z;
    }
    float4* out = reinterpret_cast<float4*>(hits + q);
    out[0] = make_float4(__uint_as_float(rec.geomID), __uint_as_float(rec.primID), rec.t, rec.u);
    out[1] = make_float4(rec.v, rec.ng[0], rec.ng[1], rec.ng[2]);
  }
}

// ------------------------------------------------------ guide photons ----
// Caustic light paths as training samples for the guide, the photon-traced
// training of parametric guiding mixtures (Vorba et al. 2014, "On-line
// learning of parametric mixture models for light transport simulation").
// Light paths leave the surface lights as Scene::sampleLightRay emits them
// (Scene.cpp:527-537) and continue through concentrated lobes only, with the
// roulette of tracePhoton (PhotonMapping.cpp:191-265).  Where one lands on a
// guided receiver with a diffuse lobe after at least one such bounce it adds
// a sample in the camera samples' layout: (x, F = luminance of the flux),
// the normal facing the photon, an outgoing direction drawn around that
// normal (a diffuse receiver's incident light does not depend on it), and
// the direction the photon came from.  Photon t of pass p uses the stream
// (seed, t, p) of its own domain.
__global__ void __launch_bounds__(kBlock) kGuidePhotons(DevScene s, uint32_t pass, uint32_t count, uint2* overflow,
                                                       float4* samples, uint32_t* sampleCount, uint32_t cap) {
  __shared__ uint2 stack[kStackLds * kBlock];
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  uint2* ovf = overflow + (size_t)gtid * kStackOvf;
  const float minQ = 1.0f - hm::powf_pos(0.1f, 1.0f / (float)s.maxDepth);
  for (uint32_t t = gtid; t < count; t += gridDim.x * blockDim.x) {
    hrng::Stream rng = hrng::make_stream(s.seed, t, pass, hrng::kDomainGuidePhoton);
    const uint32_t li = lightCdfSearch(s, rng.canonical_double());
    const int32_t ltri = (int32_t)ldg(s.lightTris + li);
    const float* pp = s.triPos + (size_t)ltri * 9;
    const V3 p0 = load3g(pp), p1 = load3g(pp + 3), p2 = load3g(pp + 6);
    const V3 ngL = normalize(cross(p1 - p0, p2 - p0));
    const float su = rng.uniform01f();
    const float tt = hm::fsqrt(rng.uniform01f());
    const float b0 = 1.0f - tt, b2 = tt * su, b1 = fmaxf(0.0f, 1.0f - b0 - b2);
    const Interaction itL = makeInteractionB(s, ltri, b0, b1, b2, ngL);
    float pDir;
    V3 rd = sampleCosHemisphere(rng, itL.ng, &pDir);
    const DevMaterial lm = ldgStruct(s.mats + itL.mat);
    const V3 Le = v3(lm.emission[0], lm.emission[1], lm.emission[2]) * (float)kInvPi;
    V3 flux = Le * (__builtin_fabsf(dot(itL.ng, rd)) * s.totalLightArea / pDir);
    V3 ro = offsetOrigin(itL, true);
    Hit h = traceClosest<kStackLds, kStackOvf>(s, ro, rd, stack + threadIdx.x, kBlock, ovf, nullptr, nullptr);
    if (h.tri < 0) continue;
    Media md;
    md.n = 0;
    md.overflow = false;
    md.tx0 = 0u;
    int wavelength = -1;
    for (int depth = 0; depth < s.maxDepth; ++depth) {
      if (!(__builtin_isfinite(flux.x) && __builtin_isfinite(flux.y) && __builtin_isfinite(flux.z))) break;
      Interaction it = makeInteraction(s, h.tri, h.u, h.v);
      DevMaterial m = ldgStruct(s.mats + it.mat);
      if (m.isLight) break;
      uint32_t texel = 0u;
      if (m.textured) applyTextures(s, m, it, 1.0f - h.u - h.v, h.u, h.v, &texel);
      const V3 wo = -rd;
      if (depth > 0 && m.hasDiffuseLobe && guideEligible(m)) {
        const V3 nf = dot(it.ng, wo) >= 0.0f ? it.ng : -it.ng;
        float pd;
        const V3 wos = sampleCosHemisphere(rng, nf, &pd);
        const float F = 0.2126f * flux.x + 0.7152f * flux.y + 0.0722f * flux.z;
        if (F > 0.0f && F < 3.0e38f) {
          const uint32_t k = atomicAdd(sampleCount, 1u);
          if (k < cap) {
            const float pB = bsdfPdf(m, it, wos, wo, mediaOutsideIOR(s, md, it, wos, wavelength));
            samples[4 * (size_t)k] = f4(it.x, F);
            samples[4 * (size_t)k + 1] = f4(nf, 0.0f);
            samples[4 * (size_t)k + 2] = f4(wos, 0.0f);
            samples[4 * (size_t)k + 3] = f4(wo, pB);
          }
        }
      }
      float sp = 1.0f;
      if (wavelength == -1 && m.varyIOR) {
        const float q0 = 1.0f / 3.0f;
        const float xi = rng.uniform01f();
        wavelength = xi < q0 ? 0 : (xi < q0 + q0 ? 1 : 2);
        sp = q0;
      }
      const float outsideIOR = mediaOutsideIOR(s, md, it, wo, wavelength);
      BsdfSample bs = bsdfSample(m, rng, it, wo, wavelength, outsideIOR, true, 2);
      V3 thr = (bs.tD + bs.tC) / sp;
      if (isZero(thr)) break;
      const V3 o2 = offsetOrigin(it, bs.outside);
      const Hit h2 = traceClosest<kStackLds, kStackOvf>(s, o2, bs.dir, stack + threadIdx.x, kBlock, ovf, nullptr,
                                                        nullptr);
      if (h2.tri < 0 || h2.tri == it.tri) break;
      mediaUpdate(md, it, wo, bs.dir, texel, nullptr);
      const Interaction it2 = makeInteraction(s, h2.tri, h2.u, h2.v);
      thr = mul(thr, beersLaw(mediaTransmittance(s, md), norm(it2.x - o2)));
      const float rq = fmaxf(minQ, 1.0f - norm(thr));
      if (rng.uniform01f() <= rq) break;
      flux = mul(flux, thr) / (1.0f - rq);
      h = h2;
      rd = bs.dir;
    }
  }
}

}  // namespace

void tileOrder(uint32_t width, uint32_t height, uint32_t* order, uint32_t* rankOf) {
  uint32_t n = 0;
  for (uint32_t ty = 0; ty < height; ty += 8)
    for (uint32_t tx = 0; tx < width; tx += 8)
      for (uint32_t y = ty; y < ty + 8 && y < height; ++y)
        for (uint32_t x = tx; x < tx + 8 && x < width; ++x) {
          order[n] = y * width + x;
          rankOf[y * width + x] = n;
          ++n;
        }
}

namespace {

void hipOk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// The bounces of one chunk: fill, extend, shade, connect, retire, advance.
// `chunk` is even, so the queue pointers are back in place afterwards and
// every chunk issues identical launches (what a captured graph replays).
template <bool MULTI, bool GUIDED, int W>
void launchChunk(const JobSet& js, const PathBuffers& pbIn, const LaunchConfig& cfg, uint32_t chunk,
                 hipStream_t stream, KernelTimer* timer) {
  PathBuffers pb = pbIn;
  auto t0 = [&](int k) {
    if (timer) timer->begin(k, stream);
  };
  auto t1 = [&](int k) {
    if (timer) timer->end(k, stream);
  };
  const int countStats = cfg.countStats ? 1 : 0;
  for (uint32_t it = 0; it < chunk; ++it) {
    t0(kKGenerate);
    kFill<MULTI><<<kFillBlocks, kBlock, 0, stream>>>(js, pb);
    t1(kKGenerate);
    t0(kKMisc);
    kFillCommit<<<1, 64, 0, stream>>>(pb.counters);
    t1(kKMisc);
    t0(kKExtend);
    if (countStats) kExtend<MULTI, true, W><<<cfg.traceBlocks, kBlock, 0, stream>>>(js, pb);
    else kExtend<MULTI, false, W><<<cfg.traceBlocks, kBlock, 0, stream>>>(js, pb);
    t1(kKExtend);
    if (GUIDED) {
      t0(kKGuideInfer);
      kGuideInfer<<<kGuideInferBlocks, kGuideInferThreads, 0, stream>>>(js, pb);
      t1(kKGuideInfer);
    }
    t0(kKShade);
    kShade<MULTI, GUIDED><<<cfg.shadeBlocks, kBlock, 0, stream>>>(js, pb);
    t1(kKShade);
    t0(kKConnect);
    if (countStats) kConnect<MULTI, true, W><<<cfg.traceBlocks, kBlock, 0, stream>>>(js, pb);
    else kConnect<MULTI, false, W><<<cfg.traceBlocks, kBlock, 0, stream>>>(js, pb);
    t1(kKConnect);
    t0(kKRetire);
    kRetire<GUIDED><<<cfg.shadeBlocks, kBlock, 0, stream>>>(js, pb, countStats ? 1u : 0u);
    t1(kKRetire);
    t0(kKMisc);
    kAdvance<<<1, 64, 0, stream>>>(pb.counters);
    t1(kKMisc);
    uint32_t* tmp = pb.qA;
    pb.qA = pb.qB;
    pb.qB = tmp;
  }
}

// The chunk's kernels for a batch (each job's node width read by the
// kernels) or for one job (its width compiled in).
void launchChunkAny(const JobSet& js, const PathBuffers& pb, const LaunchConfig& cfg, uint32_t chunk,
                    hipStream_t stream, KernelTimer* timer, bool multi, bool guided) {
  if (multi) launchChunk<true, false, 0>(js, pb, cfg, chunk, stream, timer);
  else if (guided && js.s.bvhWidth == 8) launchChunk<false, true, 8>(js, pb, cfg, chunk, stream, timer);
  else if (guided) launchChunk<false, true, 4>(js, pb, cfg, chunk, stream, timer);
  else if (js.s.bvhWidth == 8) launchChunk<false, false, 8>(js, pb, cfg, chunk, stream, timer);
  else launchChunk<false, false, 4>(js, pb, cfg, chunk, stream, timer);
}

}  // namespace

size_t jobTableBytes(uint32_t nJobs) {
  const size_t a = (sizeof(DevScene) * nJobs + 255) & ~(size_t)255;
  return a + sizeof(PassInfo) * nJobs;
}

void renderJobs(const std::vector<RenderJobDesc>& jobs, const PathBuffers& pb, const LaunchConfig& cfg,
                hipStream_t stream, KernelTimer* timer, const GuideRun* guide, void* jobTables) {
  const uint32_t nJobs = (uint32_t)jobs.size();
  if (nJobs == 0) return;
  std::vector<DevScene> scenes(nJobs);
  std::vector<PassInfo> passes(nJobs);
  uint64_t total = 0;
  int maxDepth = 0;
  for (uint32_t j = 0; j < nJobs; ++j) {
    const RenderJobDesc& d = jobs[j];
    scenes[j] = d.scene;
    PassInfo& pi = passes[j];
    pi.first = d.first;
    pi.pixels = d.scene.width * d.scene.height;
    pi.paths = d.n * pi.pixels;
    pi.pidBase = (uint32_t)total;
    pi.tileOrder = d.dTileOrder;
    pi.rankOf = d.dRankOf;
    pi.offsets = d.dOffsets;
    if (pi.pixels > pb.ringPaths) throw std::runtime_error("renderJobs: results ring smaller than one sample");
    total += (uint64_t)d.n * pi.pixels;
    maxDepth = std::max(maxDepth, d.scene.maxDepth);
  }
  if (total >= (1ull << 31)) throw std::runtime_error("renderJobs: more than 2^31 paths in one call");
  JobSet js{};
  js.s = scenes[0];
  js.pi = passes[0];
  js.nJobs = nJobs;
  js.paths = (uint32_t)total;
  if (guide) {
    if (nJobs != 1) throw std::runtime_error("renderJobs: guided rendering takes one job");
    js.g = *guide;
  }
  if (nJobs > 1) {
    if (!jobTables) throw std::runtime_error("renderJobs: no job table buffer for a batch");
    char* base = (char*)jobTables;
    void* dScenes = base;
    void* dPasses = base + ((sizeof(DevScene) * nJobs + 255) & ~(size_t)255);
    hipOk(hipMemcpyAsync(dScenes, scenes.data(), sizeof(DevScene) * nJobs, hipMemcpyHostToDevice, stream), "copy");
    hipOk(hipMemcpyAsync(dPasses, passes.data(), sizeof(PassInfo) * nJobs, hipMemcpyHostToDevice, stream), "copy");
    js.scenes = (const DevScene*)dScenes;
    js.passes = (const PassInfo*)dPasses;
  }
  if (total == 0) return;
  auto t0 = [&](int k) {
    if (timer) timer->begin(k, stream);
  };
  auto t1 = [&](int k) {
    if (timer) timer->end(k, stream);
  };
  const uint32_t chunk = std::max<uint32_t>(2, (cfg.chunk + 1) & ~1u);
  // Paths started during chunk c have finished `lag` chunks later: a path
  // has at most maxDepth vertices, one per bounce.
  const uint32_t lag = ((uint32_t)maxDepth + chunk - 1) / chunk;
  std::vector<uint32_t> acc(nJobs, 0);  // samples of each job added to its films
  uint32_t front = 0;                    // first job not completely added
  auto accPid = [&]() -> uint64_t {
    return front < nJobs ? (uint64_t)passes[front].pidBase + (uint64_t)acc[front] * passes[front].pixels : total;
  };
  uint32_t limit = (uint32_t)std::min<uint64_t>(total, accPid() + pb.ringPaths);
  t0(kKMisc);
  kInitRender<<<256, kBlock, 0, stream>>>(pb, limit);
  t1(kKMisc);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  struct GraphFree {
    hipGraph_t& g;
    hipGraphExec_t& e;
    ~GraphFree() {
      if (e) (void)hipGraphExecDestroy(e);
      if (g) (void)hipGraphDestroy(g);
    }
  } graphFree{graph, exec};
  // Per-kernel event timing needs eager launches (HIP gives no elapsed time
  // for events recorded inside a captured graph), so a timed render does not
  // replay graphs.
  if (cfg.useGraph && !timer) {
    hipOk(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    launchChunkAny(js, pb, cfg, chunk, stream, nullptr, nJobs > 1, guide != nullptr);
    hipOk(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
    hipOk(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  }
  std::vector<uint32_t> startedBy;
  for (uint32_t c = 0;; ++c) {
    if (exec) {
      t0(kKGraph);
      hipOk(hipGraphLaunch(exec, stream), "hipGraphLaunch");
      t1(kKGraph);
    } else {
      launchChunkAny(js, pb, cfg, chunk, stream, timer, nJobs > 1, guide != nullptr);
    }
    uint32_t cnt[kNumCounters];
    hipOk(hipMemcpyAsync(cnt, pb.counters, sizeof(cnt), hipMemcpyDeviceToHost, stream), "counter readback");
    hipOk(hipStreamSynchronize(stream), "hipStreamSynchronize");
    uint32_t live = cnt[kCntA], next = cnt[kCntNext];
    startedBy.push_back(next);
    uint64_t done = live == 0 ? next : (c >= lag ? startedBy[c - lag] : 0u);
    for (uint32_t j = front; j < nJobs && passes[j].pidBase < done; ++j) {
      const PassInfo& pi = passes[j];
      uint32_t n = pi.paths / pi.pixels;
      uint32_t complete = (uint32_t)std::min<uint64_t>(n, (done - pi.pidBase) / pi.pixels);
      if (complete > acc[j]) {
        uint32_t blocks = (pi.pixels + kBlock - 1) / kBlock;
        t0(kKAccumulate);
        kAccumulate<<<blocks, kBlock, 0, stream>>>(pb, pi, acc[j], complete, jobs[j].dColor, jobs[j].dNormal,
                                                   jobs[j].dAlbedo);
        t1(kKAccumulate);
        acc[j] = complete;
      }
    }
    while (front < nJobs && acc[front] * (uint64_t)passes[front].pixels == passes[front].paths) ++front;
    uint32_t newLimit = (uint32_t)std::min<uint64_t>(total, accPid() + pb.ringPaths);
    if (newLimit != limit) {
      t0(kKMisc);
      kSetLimit<<<1, 64, 0, stream>>>(pb.counters, newLimit);
      t1(kKMisc);
      limit = newLimit;
    }
    if (live == 0 && next >= total) break;
  }
  hipOk(hipStreamSynchronize(stream), "hipStreamSynchronize");
  if (front != nJobs) throw std::runtime_error("renderJobs: not every sample was accumulated");
}

void launchTraceHits(const DevScene& s, const float* dRays, uint32_t n, HitRecord* dHits, uint2* dOverflow,
                     uint32_t overflowThreads, hipStream_t stream) {
  // the grid-stride index q + stride stays below 2^32
  if (n >= 0x80000000u) throw std::runtime_error("more than 2^31 - 1 rays in one query");
  uint32_t blocks = (n + kBlock - 1) / kBlock;
  uint32_t maxBlocks = overflowThreads / kBlock;
  if (blocks > maxBlocks) blocks = maxBlocks;
  if (blocks == 0) return;
  kTraceHits<<<blocks, kBlock, 0, stream>>>(s, dRays, n, dHits, dOverflow);
}

void launchGuidePhotons(const DevScene& s, uint32_t pass, uint32_t count, uint2* dOverflow, uint32_t overflowThreads,
                        float4* samples, uint32_t* sampleCount, uint32_t cap, hipStream_t stream) {
  if (count == 0 || !s.hasSurfaceLight || s.numLights == 0) return;
  uint32_t blocks = (count + kBlock - 1) / kBlock;
  const uint32_t maxBlocks = overflowThreads / kBlock;
  if (blocks > maxBlocks) blocks = maxBlocks;
  if (blocks == 0) return;
  kGuidePhotons<<<blocks, kBlock, 0, stream>>>(s, pass, count, dOverflow, samples, sampleCount, cap);
}

void launchTraceClosest(const DevScene& s, const float* dRays, uint32_t n, int32_t* dTri, float* dTuv, uint2* dOverflow,
                        uint32_t overflowThreads, hipStream_t stream) {
  if (n >= 0x80000000u) throw std::runtime_error("more than 2^31 - 1 rays in one query");
  uint32_t blocks = (n + kBlock - 1) / kBlock;
  uint32_t maxBlocks = overflowThreads / kBlock;
  if (blocks > maxBlocks) blocks = maxBlocks;
  if (blocks == 0) return;
  kTraceRays<<<blocks, kBlock, 0, stream>>>(s, dRays, n, dTri, dTuv, dOverflow);
}

}  // namespace hasty

// C ABI (include/hasty_amd.h): scene loading and upload, BVH build, device
// memory and stream ownership, render and trace entry points, per-kernel
// HIP-event timing.
#include "../../include/hasty_amd.h"

#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "bvh.h"
#include "capi_internal.h"
#include "pt_kernels.h"
#include "rng.h"
#include "image_io.h"
#include "scene_io.h"
#include "sppm.h"
#include "guide_kernels.h"

namespace {

thread_local std::string g_lastError;

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHECK(x) check((x), #x)

// Grow-only device buffer owned by a scene (freed with it).  Callers
// synchronise their stream before a call that may grow it again.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void* ensure(size_t n) {
    if (n > bytes) {
      if (p) (void)hipFree(p);
      p = nullptr;
      bytes = 0;
      HIPCHECK(hipMalloc(&p, n));
      bytes = n;
    }
    return p;
  }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Reference Halton point generator (core/include/Hasty/Random.h:14-63).
struct Halton {
  uint32_t b, n = 0, d = 1;
  explicit Halton(uint32_t base) : b(base) {}
  float next() {
    uint32_t x = d - n;
    if (x == 1) {
      n = 1;
      d *= b;
    } else {
      uint32_t y = d / b;
      while (x <= y) y /= b;
      n = (b + 1) * y - x;
    }
    return (float)n / (float)d;
  }
};

// Offset of sample / pass index k: the (k+1)-th Halton(2, 3) point
// (RenderJob.h:53-60).
void haltonPoint(uint32_t k, float* x, float* y) {
  Halton h2(2), h3(3);
  float a = 0.0f, b = 0.0f;
  for (uint32_t i = 0; i <= k; ++i) {
    a = h2.next();
    b = h3.next();
  }
  *x = a;
  *y = b;
}

struct EventTimer : hasty::KernelTimer {
  struct Rec {
    int kernel;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::vector<hipEvent_t> open;  // one open start event per kernel id
  uint64_t launches[hasty::kNumKernels] = {0};
  double ms[hasty::kNumKernels] = {0};
  EventTimer() : open(hasty::kNumKernels, nullptr) {}
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
  }
  int only = -1;  // >= 0: time only this kernel id
  void begin(int k, hipStream_t s) override {
    if (only >= 0 && k != only) return;
    hipEvent_t e = get();
    HIPCHECK(hipEventRecord(e, s));
    open[k] = e;
  }
  void end(int k, hipStream_t s) override {
    if (only >= 0 && k != only) return;
    hipEvent_t e = get();
    HIPCHECK(hipEventRecord(e, s));
    pending.push_back({k, open[k], e});
    open[k] = nullptr;
  }
  void collect() {
    for (Rec& r : pending) {
      HIPCHECK(hipEventSynchronize(r.b));
      float t = 0.0f;
      HIPCHECK(hipEventElapsedTime(&t, r.a, r.b));
      launches[r.kernel] += 1;
      ms[r.kernel] += t;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }
  void reset() {
    collect();
    for (int k = 0; k < hasty::kNumKernels; ++k) {
      launches[k] = 0;
      ms[k] = 0.0;
    }
  }
  ~EventTimer() override {
    for (Rec& r : pending) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace

struct HastyScene;
struct BatchItem {
  HastyScene* scene;
  uint32_t seed, first, n;
  float* color;
  float* normal;
  float* albedo;
};
static void renderBatch(HastyScene* owner, const BatchItem* items, uint32_t nItems);

// A BVH of one node width resident on the scene's device.
struct DeviceBvh {
  int width = 0;
  void* nodes = nullptr;
  void* tris = nullptr;
  uint32_t numNodes = 0;
  int depth = 0;
  uint64_t bytes = 0;
};

struct HastyScene {
  hasty::Job job;
  std::vector<DeviceBvh> bvhs;  // built so far, one per width
  int bvhIndex = 0;             // the one in use (dev.nodes / dev.tris)
  hdev::DevScene dev{};
  std::vector<void*> sceneAllocs;
  uint64_t sceneBytes = 0;
  hipStream_t stream = nullptr;
  int device = 0;

  // path buffers
  hasty::PathBuffers pb{};
  std::vector<void*> pathAllocs;
  uint32_t pathCapacity = 0;
  uint32_t slotsWanted = 0;  // 0 = automatic
  uint32_t* dTileOrder = nullptr;
  uint32_t tileW = 0, tileH = 0;
  hasty::LaunchConfig launch;

  // films for the host-pointer entry point
  float* dFilm = nullptr;
  size_t filmFloats = 0;

  // trace entry point
  uint2* traceOverflow = nullptr;
  uint32_t traceOverflowThreads = 0;
  DevBuf traceIO;  // rays in, hits out of the host-pointer trace call

  // Halton offsets of a render call and the job tables of a batch
  DevBuf offsets;
  DevBuf jobTables;

  // guided rendering
  HastyGuide* guide = nullptr;
  float guideFraction = 0.5f;
  float guideLr = 0.0f;
  uint32_t guideMinibatch = 16384;
  float guideWeightCap = 20.0f;
  uint32_t guideMaxDepth = 255;    // vertices of depth below this are guided
  float guideMaxThroughput = 3.0e38f;  // ... while the path throughput is at most this
  uint32_t guidePhotons = 0;       // caustic light paths traced per training step
  float guidePhotonShare = 0.5f;   // their samples' weight sum / the camera samples'
  uint32_t guidePhotonPass = 0;    // light-path pass counter (random streams)
  uint32_t lastGuidePhotonSamples = 0;
  float4* guideRaw = nullptr;
  float4* guideRec = nullptr;
  float* guideSel = nullptr;  // per slot: learned selection probability
  float4* guideSamples = nullptr;
  uint32_t* guideCount = nullptr;
  uint32_t guideSlots = 0;
  static constexpr uint32_t kGuideSampleCap = 1u << 21;
  static constexpr uint64_t kGuideTrainPaths = 1u << 20;  // paths per online training step
  float lastGuideLoss = 0.0f;
  uint32_t lastGuideSamples = 0;

  bool profiling = false;  // HIP-event timing of every launch
  bool counting = false;   // device ray / node / triangle counters
  std::unique_ptr<EventTimer> timer;
  uint64_t stats[hasty::kNumStats] = {0};

  template <class T>
  T* upload(const T* host, size_t count) {
    if (count == 0) return nullptr;
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, count * sizeof(T)));
    sceneAllocs.push_back(p);
    sceneBytes += count * sizeof(T);
    HIPCHECK(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
    return (T*)p;
  }
  template <class T>
  T* allocPath(size_t count) {
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, count * sizeof(T)));
    pathAllocs.push_back(p);
    return (T*)p;
  }
  void freePath() {
    for (void* p : pathAllocs) (void)hipFree(p);
    pathAllocs.clear();
    pathCapacity = 0;
  }
  ~HastyScene() {
    if (timer) timer->collect();
    freePath();
    for (DeviceBvh& b : bvhs) {
      if (b.nodes) (void)hipFree(b.nodes);
      if (b.tris) (void)hipFree(b.tris);
    }
    for (void* p : sceneAllocs) (void)hipFree(p);
    if (dTileOrder) (void)hipFree(dTileOrder);
    for (void* p : {(void*)guideRaw, (void*)guideRec, (void*)guideSel, (void*)guideSamples, (void*)guideCount})
      if (p) (void)hipFree(p);
    if (dFilm) (void)hipFree(dFilm);
    if (traceOverflow) (void)hipFree(traceOverflow);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void applySettings() {
    const hasty::RenderSettings& st = job.settings;
    dev.maxDepth = st.maxDepth;
    dev.rrStart = st.rouletteStartDepth;
    dev.rrQ = st.rouletteQ;
    dev.useMIS = st.useMIS ? 1 : 0;
    dev.width = st.width;
    dev.height = st.height;
  }

  // Path slots in flight (regenerated as paths end) and the results ring.
  // Slots for a render of `paths` paths: the wanted pool (default 2^23), but
  // no more than the render needs (at least 2^16).
  uint32_t slotCount(uint64_t paths = 0) const {
    if (paths == 0) paths = (uint64_t)dev.width * dev.height;
    uint64_t want = slotsWanted ? slotsWanted : (1ull << 23);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, std::max<uint64_t>(paths, 1ull << 16)));
  }
  uint32_t ringPathsWanted() const {
    uint64_t P = (uint64_t)dev.width * dev.height;
    uint64_t r = std::max<uint64_t>(P, (uint64_t)(1u << 26) / P * P);  // about 64M paths, whole samples
    return (uint32_t)std::min<uint64_t>(r, 0x7FFFFFFFull / P * P);
  }

  void ensurePathBuffers(uint64_t paths = 0) {
    uint32_t need = slotCount(paths);
    uint32_t ring = ringPathsWanted();
    uint32_t traceThreads = launch.traceBlocks * hasty::kBlock;
    if (need == pathCapacity && ring == pb.ringPaths && pb.traceThreads == traceThreads) return;
    freePath();
    size_t n = need;
    pb.rayO = allocPath<float4>(n);
    pb.rayD = allocPath<float4>(n);
    pb.hit = allocPath<float4>(n);
    pb.beta = allocPath<float4>(n);
    pb.L = allocPath<float4>(n);
    pb.state = allocPath<uint4>(n);
    pb.media = allocPath<uint4>(2 * n);
    pb.mediaTx = allocPath<uint32_t>((size_t)hdev::kMaxMedia * n);
    pb.auxN = allocPath<float4>(n);
    pb.auxA = allocPath<float4>(n);
    pb.shO = allocPath<float4>(n);
    pb.shD = allocPath<float4>(n);
    pb.shFD = allocPath<float4>(n);
    pb.shFC = allocPath<float4>(n);
    pb.shThr = allocPath<float4>(n);
    pb.shBeta = allocPath<float4>(n);
    pb.shT = allocPath<float4>(n);
    pb.qA = allocPath<uint32_t>(n);
    pb.qB = allocPath<uint32_t>(n);
    pb.qS = allocPath<uint32_t>(n);
    pb.qR = allocPath<uint32_t>(n);
    pb.mark = allocPath<uint32_t>(n);
    pb.counters = allocPath<uint32_t>(hasty::kNumCounters);
    pb.stats = allocPath<unsigned long long>(hasty::kNumStats);
    pb.stackOverflow = allocPath<uint2>((size_t)traceThreads * hasty::kStackOvf);
    pb.ringL = allocPath<float4>(ring);
    pb.ringN = allocPath<float4>(ring);
    pb.ringA = allocPath<float4>(ring);
    pb.ringPaths = ring;
    pb.traceThreads = traceThreads;
    pb.capacity = need;
    pathCapacity = need;
    HIPCHECK(hipMemset(pb.stats, 0, sizeof(unsigned long long) * hasty::kNumStats));
  }

  void ensureTileOrder() {
    if (tileW == dev.width && tileH == dev.height && dTileOrder) return;
    if (dTileOrder) (void)hipFree(dTileOrder);
    size_t P = (size_t)dev.width * dev.height;
    std::vector<uint32_t> t(2 * P);
    hasty::tileOrder(dev.width, dev.height, t.data(), t.data() + P);
    HIPCHECK(hipMalloc(&dTileOrder, sizeof(uint32_t) * t.size()));
    HIPCHECK(hipMemcpy(dTileOrder, t.data(), sizeof(uint32_t) * t.size(), hipMemcpyHostToDevice));
    tileW = dev.width;
    tileH = dev.height;
  }
  const uint32_t* rankOf() const { return dTileOrder + (size_t)dev.width * dev.height; }

  const DeviceBvh& bvh() const { return bvhs[bvhIndex]; }

  // Makes the BVH of node width `width` (4 or 8) current, building and
  // uploading it the first time.  Throws if its depth does not fit the
  // traversal stack ((width - 1) * depth + 1 entries per lane).
  void useBvhWidth(int width) {
    for (size_t i = 0; i < bvhs.size(); ++i)
      if (bvhs[i].width == width) {
        bvhIndex = (int)i;
        dev.nodes = (const float4*)bvhs[i].nodes;
        dev.tris = (const float4*)bvhs[i].tris;
        dev.bvhWidth = width;
        return;
      }
    const hasty::Scene& sc = job.scene;
    hasty::BvhBuildOptions opt;
    hasty::Bvh4 b = hasty::buildBvh4(sc.triPos.data(), sc.numTris, opt, width);
    if ((width - 1) * b.depth + 1 > hasty::kStackLds + hasty::kStackOvf)
      throw std::runtime_error("BVH too deep for the traversal stack (width " + std::to_string(width) + ", depth " +
                               std::to_string(b.depth) + ")");
    if ((uint64_t)b.nodes.size() * sizeof(float) >= (1ull << 32))  // traversal uses 32-bit node offsets
      throw std::runtime_error("BVH larger than 4 GiB (" + std::to_string(b.numNodes) + " nodes)");
    DeviceBvh d;
    d.width = width;
    d.numNodes = b.numNodes;
    d.depth = b.depth;
    const size_t nb = b.nodes.size() * sizeof(float), tb = b.tris.size() * sizeof(float);
    struct Guard {
      DeviceBvh& d;
      bool keep = false;
      ~Guard() {
        if (keep) return;
        if (d.nodes) (void)hipFree(d.nodes);
        if (d.tris) (void)hipFree(d.tris);
      }
    } guard{d};
    HIPCHECK(hipMalloc(&d.nodes, std::max<size_t>(nb, 16)));
    HIPCHECK(hipMalloc(&d.tris, std::max<size_t>(tb, 16)));
    if (nb) HIPCHECK(hipMemcpy(d.nodes, b.nodes.data(), nb, hipMemcpyHostToDevice));
    if (tb) HIPCHECK(hipMemcpy(d.tris, b.tris.data(), tb, hipMemcpyHostToDevice));
    d.bytes = nb + tb;
    guard.keep = true;
    bvhs.push_back(d);
    useBvhWidth(width);
  }

  void renderDevice(uint32_t seed, uint32_t first, uint32_t nSamples, float* dC, float* dN, float* dA) {
    BatchItem it{this, seed, first, nSamples, dC, dN, dA};
    renderBatch(this, &it, 1);
  }
};

// Halton offsets of samples first .. first+n-1 (point index k+1).
static std::vector<float2> haltonOffsets(uint32_t first, uint32_t n) {
  Halton h2(2), h3(3);
  std::vector<float2> offs(n);
  for (uint32_t k = 0; k < first + n; ++k) {
    float a = h2.next(), b = h3.next();
    if (k >= first) offs[k - first] = make_float2(a, b);
  }
  return offs;
}

// Renders the items through the path pool and stream of `owner`.  Items are
// grouped so that each renderJobs call stays below 2^31 paths; a single item
// larger than that is split by samples.
static void renderGuided(HastyScene* s, const BatchItem& it);

static void renderBatch(HastyScene* owner, const BatchItem* items, uint32_t nItems) {
  if (owner->guide) {
    if (nItems != 1 || items[0].scene != owner) throw std::runtime_error("guided rendering takes one job of its scene");
    renderGuided(owner, items[0]);
    return;
  }
  uint64_t allPaths = 0;
  for (uint32_t i = 0; i < nItems; ++i)
    allPaths += (uint64_t)items[i].n * items[i].scene->dev.width * items[i].scene->dev.height;
  owner->ensurePathBuffers(allPaths);
  // Split every item into calls of < 2^31 paths and lay out all Halton
  // offsets in one upload to the owner's offset buffer.
  struct Piece {
    uint32_t item, first, n;
    size_t off;
  };
  std::vector<Piece> pieces;
  std::vector<float2> offs;
  const uint64_t maxPaths = 0x7FFFFFFFull;
  for (uint32_t i = 0; i < nItems; ++i) {
    const BatchItem& it = items[i];
    HastyScene* sc = it.scene;
    if (sc->device != owner->device) throw std::runtime_error("batch jobs must live on one device");
    if (sc->dev.width == 0 || sc->dev.height == 0) throw std::runtime_error("empty film");
    sc->ensureTileOrder();
    uint64_t P = (uint64_t)sc->dev.width * sc->dev.height;
    uint32_t perCall = (uint32_t)std::max<uint64_t>(1, maxPaths / P);
    for (uint32_t done = 0; done < it.n; done += perCall) {
      uint32_t k = std::min(perCall, it.n - done);
      pieces.push_back(Piece{i, it.first + done, k, offs.size()});
      std::vector<float2> o = haltonOffsets(it.first + done, k);
      offs.insert(offs.end(), o.begin(), o.end());
    }
  }
  float2* dOffs = (float2*)owner->offsets.ensure(sizeof(float2) * std::max<size_t>(1, offs.size()));
  if (!offs.empty())
    HIPCHECK(hipMemcpyAsync(dOffs, offs.data(), sizeof(float2) * offs.size(), hipMemcpyHostToDevice, owner->stream));
  void* tables = owner->jobTables.ensure(hasty::jobTableBytes((uint32_t)std::max<size_t>(1, pieces.size())));
  std::vector<hasty::RenderJobDesc> group;
  uint64_t groupPaths = 0;
  owner->launch.countStats = owner->counting;
  if (owner->profiling && !owner->timer) owner->timer = std::make_unique<EventTimer>();
  hasty::KernelTimer* timer = owner->profiling ? owner->timer.get() : nullptr;
  auto flush = [&]() {
    if (group.empty()) return;
    hasty::renderJobs(group, owner->pb, owner->launch, owner->stream, timer, nullptr, tables);
    HIPCHECK(hipGetLastError());
    group.clear();
    groupPaths = 0;
  };
  for (const Piece& pc : pieces) {
    const BatchItem& it = items[pc.item];
    HastyScene* sc = it.scene;
    uint64_t P = (uint64_t)sc->dev.width * sc->dev.height;
    if (groupPaths + (uint64_t)pc.n * P > maxPaths) flush();
    hasty::RenderJobDesc jd;
    jd.scene = sc->dev;
    jd.scene.seed = it.seed;
    jd.first = pc.first;
    jd.n = pc.n;
    jd.dOffsets = dOffs + pc.off;
    jd.dTileOrder = sc->dTileOrder;
    jd.dRankOf = sc->rankOf();
    jd.dColor = it.color;
    jd.dNormal = it.normal;
    jd.dAlbedo = it.albedo;
    group.push_back(jd);
    groupPaths += (uint64_t)pc.n * P;
  }
  flush();
  HIPCHECK(hipStreamSynchronize(owner->stream));
  if (owner->counting) {
    unsigned long long s[hasty::kNumStats];
    HIPCHECK(hipMemcpy(s, owner->pb.stats, sizeof(s), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemset(owner->pb.stats, 0, sizeof(s)));
    for (int i = 0; i < hasty::kNumStats; ++i) owner->stats[i] += s[i];
  }
  if (owner->profiling) owner->timer->collect();
}

struct HastyGuide {
  HastyScene* scene = nullptr;
  std::unique_ptr<guide::Guide> g;
  float4* dBuf = nullptr;
  size_t bufFloats = 0;
  float4* buffer(size_t floats) {
    if (floats > bufFloats) {
      if (dBuf) (void)hipFree(dBuf);
      dBuf = nullptr;
      if (hipMalloc(&dBuf, floats * sizeof(float)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
      bufFloats = floats;
    }
    return dBuf;
  }
  ~HastyGuide() {
    g.reset();
    if (dBuf) (void)hipFree(dBuf);
  }
};

struct HastySppm {
  HastyScene* scene = nullptr;
  std::unique_ptr<hasty::SppmRenderer> r;
  float* dFilm = nullptr;
  ~HastySppm() {
    r.reset();
    if (dFilm) (void)hipFree(dFilm);
  }
};

// Guided rendering.  Training: whole sample indices in groups of about 2^20
// paths, each group rendered with the current guide and followed by Adam
// steps over the radiance samples its paths produced: one pass of
// minibatches, and at least one step per sample index of the group.  Without
// training: every index in one call.
static void renderGuided(HastyScene* s, const BatchItem& it) {
  s->ensurePathBuffers((uint64_t)it.n * s->dev.width * s->dev.height);
  s->ensureTileOrder();
  uint32_t slots = s->pb.capacity;
  if (s->guideSlots != slots) {
    for (void* p : {(void*)s->guideRaw, (void*)s->guideRec, (void*)s->guideSel})
      if (p) (void)hipFree(p);
    s->guideRaw = nullptr, s->guideRec = nullptr, s->guideSel = nullptr;
    s->guideSlots = 0;
    HIPCHECK(hipMalloc(&s->guideRaw, sizeof(float4) * 10 * (size_t)slots));
    HIPCHECK(hipMalloc(&s->guideRec, sizeof(float4) * 12 * (size_t)slots));
    s->guideSlots = slots;
  }
  if (s->guide->g->learnedSelection() && !s->guideSel) {
    HIPCHECK(hipMalloc(&s->guideSel, sizeof(float) * (size_t)slots));
  }
  if (!s->guideSamples) {
    HIPCHECK(hipMalloc(&s->guideSamples, sizeof(float4) * 4 * (size_t)HastyScene::kGuideSampleCap));
    HIPCHECK(hipMalloc(&s->guideCount, sizeof(uint32_t)));
  }
  hasty::GuideRun gr;
  gr.net = s->guide->g->net();
  gr.raw = s->guideRaw;
  gr.rec = s->guideRec;
  gr.samples = s->guideSamples;
  gr.sampleCount = s->guideCount;
  // Without training no radiance samples are recorded.
  const bool training = s->guideLr > 0.0f;
  gr.sampleCap = training ? HastyScene::kGuideSampleCap : 0u;
  gr.fraction = s->guideFraction;
  gr.sel = s->guide->g->learnedSelection() ? s->guideSel : nullptr;
  gr.maxDepth = s->guideMaxDepth;
  gr.maxThroughput = s->guideMaxThroughput;
  s->launch.countStats = s->counting;
  if (s->profiling && !s->timer) s->timer = std::make_unique<EventTimer>();
  hasty::KernelTimer* timer = s->profiling ? s->timer.get() : nullptr;
  const uint64_t P = (uint64_t)s->dev.width * s->dev.height;
  const uint64_t groupPaths = training ? HastyScene::kGuideTrainPaths : 0x7FFFFFFFull;
  const uint32_t perCall = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(it.n, groupPaths / P));
  float2* dOff = (float2*)s->offsets.ensure(sizeof(float2) * perCall);
  for (uint32_t k = 0; k < it.n; k += perCall) {
    const uint32_t cnt = std::min(perCall, it.n - k);
    std::vector<float2> off = haltonOffsets(it.first + k, cnt);
    HIPCHECK(hipMemcpyAsync(dOff, off.data(), sizeof(float2) * cnt, hipMemcpyHostToDevice, s->stream));
    HIPCHECK(hipMemsetAsync(s->guideCount, 0, sizeof(uint32_t), s->stream));
    hasty::RenderJobDesc jd;
    jd.scene = s->dev;
    jd.scene.seed = it.seed;
    jd.first = it.first + k;
    jd.n = cnt;
    jd.dOffsets = dOff;
    jd.dTileOrder = s->dTileOrder;
    jd.dRankOf = s->rankOf();
    jd.dColor = it.color;
    jd.dNormal = it.normal;
    jd.dAlbedo = it.albedo;
    hasty::renderJobs({jd}, s->pb, s->launch, s->stream, timer, &gr);
    if (training) {
      uint32_t n = 0;
      HIPCHECK(hipMemcpy(&n, s->guideCount, sizeof(uint32_t), hipMemcpyDeviceToHost));
      n = std::min(n, HastyScene::kGuideSampleCap);
      if (s->guidePhotons > 0 && n > 0 && n < HastyScene::kGuideSampleCap && s->dev.hasSurfaceLight) {
        // caustic light paths append their samples after the camera samples,
        // scaled so that they carry guidePhotonShare of the camera weight
        const float camSum = s->guide->g->weightSum(s->guideSamples, n);
        HIPCHECK(hipStreamSynchronize(s->stream));
        HIPCHECK(hipMemcpy(s->guideCount, &n, sizeof(uint32_t), hipMemcpyHostToDevice));
        hdev::DevScene ds = s->dev;
        ds.seed = it.seed;
        hasty::launchGuidePhotons(ds, s->guidePhotonPass++, s->guidePhotons, s->pb.stackOverflow,
                                  s->pb.traceThreads, s->guideSamples, s->guideCount, HastyScene::kGuideSampleCap,
                                  s->stream);
        HIPCHECK(hipGetLastError());
        uint32_t all = 0;
        HIPCHECK(hipStreamSynchronize(s->stream));
        HIPCHECK(hipMemcpy(&all, s->guideCount, sizeof(uint32_t), hipMemcpyDeviceToHost));
        all = std::min(all, HastyScene::kGuideSampleCap);
        float4* ph = s->guideSamples + 4 * (size_t)n;
        const float phSum = s->guide->g->weightSum(ph, all - n);
        if (phSum > 0.0f && camSum > 0.0f) {
          s->guide->g->scaleWeights(ph, all - n, s->guidePhotonShare * camSum / phSum);
          s->lastGuidePhotonSamples = all - n;
          n = all;
        } else {
          s->lastGuidePhotonSamples = 0;
        }
      }
      s->lastGuideSamples = n;
      if (n > 0)
        s->lastGuideLoss =
            s->guide->g->trainEpoch(s->guideSamples, n, s->guideLr, s->guideWeightCap, s->guideMinibatch, cnt);
      gr.net = s->guide->g->net();
    }
  }
  HIPCHECK(hipStreamSynchronize(s->stream));
  if (s->counting) {
    unsigned long long st[hasty::kNumStats];
    HIPCHECK(hipMemcpy(st, s->pb.stats, sizeof(st), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemset(s->pb.stats, 0, sizeof(st)));
    for (int i = 0; i < hasty::kNumStats; ++i) s->stats[i] += st[i];
  }
  if (s->profiling) s->timer->collect();
}

// The wavefront path state keeps a path's depth in 8 bits.
static void checkSettings(const hasty::RenderSettings& st) {
  if (st.maxDepth > 255) throw std::runtime_error("max_depth above 255 is not supported");
}

// Scene::Scene's inputs from a descriptor (what tinyobj and OIIO hand it).
static void jobFromDesc(const HastySceneDesc& d, hasty::Job& job) {
  hasty::RenderSettings& st = job.settings;
  st.rouletteStartDepth = d.settings.roulette_start_depth;
  st.rouletteQ = d.settings.roulette_q;
  st.width = d.settings.width;
  st.height = d.settings.height;
  st.numThreads = d.settings.num_threads;
  st.maxDepth = d.settings.max_depth;
  st.numSamples = d.settings.num_samples;
  st.useMIS = d.settings.use_mis != 0;
  hasty::Scene& sc = job.scene;
  hasty::Camera& c = sc.camera;
  for (int i = 0; i < 3; ++i) {
    c.position[i] = d.camera.position[i];
    c.forward[i] = d.camera.forward[i];
    c.right[i] = d.camera.right[i];
    c.up[i] = d.camera.up[i];
  }
  c.fovDegree = d.camera.fov_degree;
  c.aperture = d.camera.aperture;
  c.focalDistance = d.camera.focal_distance;
  c.bladeRotation = d.camera.blade_rotation;
  c.numBlades = d.camera.num_blades;
  c.exposure = d.camera.exposure;
  if (c.aperture > 0.0f && c.numBlades < 3) throw std::runtime_error("camera num_blades must be >= 3");
  auto need = [](const void* p, uint32_t n, const char* what) {
    if (n > 0 && !p) throw std::runtime_error(std::string("descriptor: ") + what + " is null");
  };
  need(d.vertices, d.num_vertices, "vertices");
  need(d.normals, d.num_normals, "normals");
  need(d.texcoords, d.num_texcoords, "texcoords");
  need(d.materials, d.num_materials, "materials");
  for (const void* p : {(const void*)d.vertex_indices, (const void*)d.geom_ids, (const void*)d.prim_ids,
                        (const void*)d.material_ids})
    need(p, d.num_triangles, "a per-triangle array");
  sc.vertices.assign(d.vertices, d.vertices + 3 * (size_t)d.num_vertices);
  sc.normals.assign(d.normals, d.normals + 3 * (size_t)d.num_normals);
  sc.texcoords.assign(d.texcoords, d.texcoords + 2 * (size_t)d.num_texcoords);
  need(d.textures, d.num_textures, "textures");
  for (uint32_t t = 0; t < d.num_textures; ++t) {
    const HastyTextureDesc& td = d.textures[t];
    if (td.channels != 1 && td.channels != 3) throw std::runtime_error("descriptor: textures have 1 or 3 channels");
    if (td.width == 0 || td.height == 0 || !td.pixels) throw std::runtime_error("descriptor: empty texture");
    hasty::Texture tx;
    tx.path = "texture" + std::to_string(t);
    tx.width = td.width;
    tx.height = td.height;
    tx.channels = td.channels;
    tx.px.assign(td.pixels, td.pixels + (size_t)td.width * td.height * td.channels);
    sc.textures.push_back(std::move(tx));
  }
  std::vector<uint8_t> tinted;
  for (uint32_t m = 0; m < d.num_materials; ++m) {
    const HastyMaterialDesc& md = d.materials[m];
    hasty::Material out;
    out.name = "material" + std::to_string(m);
    for (int i = 0; i < 3; ++i) {
      out.albedo[i] = md.albedo[i];
      out.emission[i] = md.emission[i];
    }
    out.roughness = md.roughness;
    out.metallic = md.metallic;
    out.specular = md.specular;
    out.ior = md.ior;
    out.anisotropy = md.anisotropy;
    out.transmission = md.transmission;
    out.varyIOR = md.vary_ior != 0;
    if (d.material_textures) {
      const int32_t* t = d.material_textures + 4 * (size_t)m;
      for (int k = 0; k < 4; ++k)
        if (t[k] < -1 || t[k] >= (int32_t)d.num_textures) throw std::runtime_error("descriptor: texture index out of range");
      out.texAlbedo = t[0];
      out.texRoughness = t[1];
      out.texMetallic = t[2];
      out.texNormal = t[3];
    }
    sc.materials.push_back(out);
    tinted.push_back(md.tinted_emitter != 0);
  }
  // shapes from (geom_id, prim_id): shape by shape, faces in order
  std::vector<hasty::ShapeFaces> shapes;
  for (uint32_t i = 0; i < d.num_triangles; ++i) {
    const uint32_t g = d.geom_ids[i];
    if (!shapes.empty() && g < shapes.size() - 1) throw std::runtime_error("descriptor: geom_ids decrease");
    if (g >= (1u << 16)) throw std::runtime_error("more than 65535 shapes");
    while (shapes.size() <= g) {
      shapes.emplace_back();
      shapes.back().name = "shape" + std::to_string(shapes.size() - 1);
    }
    hasty::ShapeFaces& sh = shapes[g];
    if (d.prim_ids[i] != sh.materialIds.size()) throw std::runtime_error("descriptor: prim_ids must count faces from 0");
    for (int v = 0; v < 3; ++v) {
      hasty::FaceIndex ix;
      ix.v = d.vertex_indices[3 * (size_t)i + v];
      ix.n = d.normal_indices ? d.normal_indices[3 * (size_t)i + v] : -1;
      ix.t = d.texcoord_indices ? d.texcoord_indices[3 * (size_t)i + v] : -1;
      sh.indices.push_back(ix);
    }
    sh.materialIds.push_back(d.material_ids[i]);
    sh.smoothing.push_back(d.smooth && d.smooth[i] ? 1u : 0u);
  }
  if (d.has_environment) {
    need(d.env_rgb, d.env_width * d.env_height, "env_rgb");
    hasty::ImageRGB img;
    img.width = d.env_width;
    img.height = d.env_height;
    img.rgb.assign(d.env_rgb, d.env_rgb + (size_t)d.env_width * d.env_height * 3);
    hasty::setEnvironment(sc, std::move(img));
  } else {
    sc.hasEnvTexture = false;
    for (int i = 0; i < 3; ++i) sc.backgroundColor[i] = d.background[i];
  }
  hasty::buildTriangles(sc, shapes, tinted);
}

static void uploadScene(HastyScene& S);

static void buildScene(HastyScene& S, const char* path) {
  hasty::loadJob(path, S.job);
  uploadScene(S);
}

static void uploadScene(HastyScene& S) {
  hasty::Scene& sc = S.job.scene;
  const hasty::RenderSettings& st = S.job.settings;
  checkSettings(st);
  if (sc.shapeNames.size() >= 65536) throw std::runtime_error("more than 65535 shapes");
  if (sc.materials.size() >= 65536) throw std::runtime_error("more than 65535 materials");
  std::vector<hdev::DevMaterial> mats(sc.materials.size());
  for (size_t i = 0; i < sc.materials.size(); ++i) {
    const hasty::Material& m = sc.materials[i];
    hdev::DevMaterial& d = mats[i];
    for (int c = 0; c < 3; ++c) {
      d.albedo[c] = m.albedo[c];
      d.emission[c] = m.emission[c];
    }
    d.roughness = m.roughness;
    float alpha = m.roughness * m.roughness;  // PrincipledBRDF::computeAlpha
    if (alpha < 1e-2f && alpha != 0.0f) alpha = 1e-2f;
    d.alpha = alpha;
    d.metallic = m.metallic;
    d.specular = m.specular;
    d.ior = m.ior;
    d.anisotropy = m.anisotropy;
    d.transmission = m.transmission;
    d.varyIOR = m.varyIOR ? 1 : 0;
    d.isLight = m.isLight() ? 1 : 0;
    d.hasDiffuseLobe = ((1.0f - 0.5f * m.specular) * (1.0 - m.transmission) * (1.0 - m.metallic) > 0.0f) ? 1 : 0;
    d.tex[0] = m.texAlbedo;
    d.tex[1] = m.texRoughness;
    d.tex[2] = m.texMetallic;
    d.tex[3] = m.texNormal;
    d.textured = m.textured() ? 1 : 0;
    for (int t : {m.texAlbedo, m.texRoughness, m.texMetallic, m.texNormal})
      if (t >= (int)sc.textures.size()) throw std::runtime_error("texture index out of range");
    if (m.texAlbedo >= 0 && sc.textures[m.texAlbedo].channels != 3) throw std::runtime_error("albedo textures have 3 channels");
    if (m.texNormal >= 0 && sc.textures[m.texNormal].channels != 3) throw std::runtime_error("normal maps have 3 channels");
    if (m.texRoughness >= 0 && sc.textures[m.texRoughness].channels != 1)
      throw std::runtime_error("roughness textures have 1 channel");
    if (m.texMetallic >= 0 && sc.textures[m.texMetallic].channels != 1)
      throw std::runtime_error("metallic textures have 1 channel");
    // Rough transmission throws "not implemented" when shaded (BRDF.cpp:386-390);
    // the GPU path rejects such a scene when it is loaded.
    if (m.transmission > 0.0f && alpha != 0.0f && m.texRoughness < 0) throw std::runtime_error("not implemented");
    if (m.transmission > 0.0f && m.texRoughness >= 0) {
      // a roughness texture on a transmissive material: every texel must give
      // alpha 0, and every face must have uv (roughness -1 without)
      for (float r : sc.textures[m.texRoughness].px) {
        float a = r * r;
        if (a != 0.0f) throw std::runtime_error("not implemented");
      }
      for (uint32_t t = 0; t < sc.numTris; ++t)
        if (sc.triMaterial[t] == (int32_t)i && !sc.triHasUV[t]) throw std::runtime_error("not implemented");
    }
  }
  // Scene.cpp:508-512 exits the process when NEE needs a light triangle and
  // there is none; reject the scene instead.
  if (sc.hasLight() && st.useMIS && !sc.hasEnvTexture && sc.lightTris.empty())
    throw std::runtime_error("error: no light triangle found ");


  HIPCHECK(hipGetDevice(&S.device));
  HIPCHECK(hipStreamCreate(&S.stream));
  hdev::DevScene& d = S.dev;
  S.useBvhWidth(4);
  d.triPos = S.upload(sc.triPos.data(), sc.triPos.size());
  d.triNrm = S.upload(sc.triNrm.data(), sc.triNrm.size());
  d.triMat = S.upload(sc.triMaterial.data(), sc.triMaterial.size());
  d.triGeom = S.upload(sc.triGeomID.data(), sc.triGeomID.size());
  d.shapeTriOffset = S.upload(sc.shapeTriOffset.data(), sc.shapeTriOffset.size());
  d.triSmooth = S.upload(sc.triSmooth.data(), sc.triSmooth.size());
  d.mats = S.upload(mats.data(), mats.size());
  d.texturedMedia = 0;
  bool anyTexture = false;
  for (const hasty::Material& m : sc.materials) {
    anyTexture = anyTexture || m.textured();
    if (m.texAlbedo >= 0) d.texturedMedia = 1;
  }
  if (anyTexture) {
    std::vector<hdev::DevTexture> texs(sc.textures.size());
    for (size_t t = 0; t < sc.textures.size(); ++t) {
      const hasty::Texture& tx = sc.textures[t];
      if (tx.width == 0 || tx.height == 0 || tx.px.size() != (size_t)tx.width * tx.height * tx.channels)
        throw std::runtime_error("texture " + tx.path + " has no pixels");
      if ((uint64_t)tx.width * tx.height >= 0xFFFFFFFFull) throw std::runtime_error("texture " + tx.path + " too large");
      texs[t].px = S.upload(tx.px.data(), tx.px.size());
      texs[t].w = tx.width;
      texs[t].h = tx.height;
    }
    d.texs = S.upload(texs.data(), texs.size());
    d.triUV = S.upload(sc.triUV.data(), sc.triUV.size());
    d.triHasUV = S.upload(sc.triHasUV.data(), sc.triHasUV.size());
  }
  d.lightTris = S.upload(sc.lightTris.data(), sc.lightTris.size());
  d.lightCdf = S.upload(sc.lightCdf.data(), sc.lightCdf.size());
  d.numLights = (uint32_t)sc.lightTris.size();
  d.totalLightArea = sc.totalLightArea;
  d.hasEnv = sc.hasEnvTexture ? 1 : 0;
  if (sc.hasEnvTexture) {
    d.envRGB = S.upload(sc.env.rgb.data(), sc.env.rgb.size());
    d.envPdf = S.upload(sc.envPdf.data(), sc.envPdf.size());
    d.envCdf = S.upload(sc.envCdf.data(), sc.envCdf.size());
    std::vector<uint32_t> guide(hdev::kEnvGuide + 1);
    uint32_t n = (uint32_t)sc.envCdf.size();
    for (uint32_t g = 0; g <= hdev::kEnvGuide; ++g)
      guide[g] = hrng::upper_bound(sc.envCdf.data(), n, g / (double)hdev::kEnvGuide);
    d.envCdfGuide = S.upload(guide.data(), guide.size());
    d.envW = sc.env.width;
    d.envH = sc.env.height;
    d.envValueSum = sc.envValueSum;
  }
  for (int c = 0; c < 3; ++c) {
    d.bg[c] = sc.backgroundColor[c];
    d.camPos[c] = sc.camera.position[c];
    d.camFwd[c] = sc.camera.forward[c];
    d.camRight[c] = sc.camera.right[c];
    d.camUp[c] = sc.camera.up[c];
  }
  d.hasLight = sc.hasLight() ? 1 : 0;
  d.hasSurfaceLight = sc.hasSurfaceLight() ? 1 : 0;
  d.fovDegree = sc.camera.fovDegree;
  d.aperture = sc.camera.aperture;
  d.focalDistance = sc.camera.focalDistance;
  d.bladeRotation = sc.camera.bladeRotation;
  d.numBlades = sc.camera.numBlades;
  d.tanHalfFov = hdev::tanHalfFov(d.fovDegree);
  d.ngonNx = d.numBlades >= 3 ? hm::cosf(hdev::ngonSector(d.numBlades) * 0.5f) : 0.0f;
  d.ngonNy = d.numBlades >= 3 ? hm::sinf(hdev::ngonSector(d.numBlades) * 0.5f) : 0.0f;
  S.applySettings();
}

// The scene's traversal stack overflow buffer (256 blocks of threads).
static void ensureTraceOverflow(HastyScene* s) {
  if (!s->traceOverflow) {
    const uint32_t threads = 256 * hasty::kBlock;
    HIPCHECK(hipMalloc(&s->traceOverflow, sizeof(uint2) * threads * hasty::kStackOvf));
    s->traceOverflowThreads = threads;
  }
}

namespace hasty {
bool sceneView(const HastyScene* s, SceneView* out) {
  if (!s || !out) return false;
  out->dev = s->dev;
  out->stream = s->stream;
  out->device = s->device;
  out->bvhDepth = s->bvh().depth;
  out->bvhWidth = s->bvh().width;
  return true;
}
void setLastError(const std::string& message) { g_lastError = message; }
}  // namespace hasty

static int fail(const std::exception& e, int code) {
  g_lastError = e.what();
  return code;
}

extern "C" {

const char* hasty_last_error(void) { return g_lastError.c_str(); }

int hasty_image_read(const char* path, uint32_t* width, uint32_t* height, uint32_t* channels, float* pixels,
                     uint64_t capacity) {
  if (!path || !width || !height || !channels) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    hasty::FloatImage im = hasty::readImage(path);
    *width = im.width;
    *height = im.height;
    *channels = im.channels;
    if (pixels) {
      if (capacity < im.px.size()) throw std::runtime_error("pixel buffer too small");
      memcpy(pixels, im.px.data(), im.px.size() * sizeof(float));
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 2);
  }
}

int hasty_image_write_exr(const char* path, uint32_t width, uint32_t height, const float* rgb) {
  if (!path || (!rgb && (uint64_t)width * height > 0)) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    hasty::writeEXR(path, width, height, rgb);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 2);
  }
}

int hasty_scene_load(const char* path, HastyScene** out) {
  if (!path || !out) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    auto S = std::make_unique<HastyScene>();
    buildScene(*S, path);
    *out = S.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 2);
  }
}

int hasty_scene_create(const HastySceneDesc* desc, HastyScene** out) {
  if (!desc || !out) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    auto S = std::make_unique<HastyScene>();
    jobFromDesc(*desc, S->job);
    uploadScene(*S);
    *out = S.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 2);
  }
}

void hasty_scene_destroy(HastyScene* s) { delete s; }

int hasty_scene_info(const HastyScene* s, HastySceneInfo* info) {
  if (!s || !info) {
    g_lastError = "null argument";
    return 1;
  }
  const hasty::RenderSettings& st = s->job.settings;
  info->settings.roulette_start_depth = st.rouletteStartDepth;
  info->settings.roulette_q = st.rouletteQ;
  info->settings.width = st.width;
  info->settings.height = st.height;
  info->settings.num_threads = st.numThreads;
  info->settings.max_depth = st.maxDepth;
  info->settings.num_samples = st.numSamples;
  info->settings.use_mis = st.useMIS ? 1 : 0;
  info->num_triangles = s->job.scene.numTris;
  info->num_shapes = (uint32_t)s->job.scene.shapeNames.size();
  info->num_materials = (uint32_t)s->job.scene.materials.size();
  info->num_light_triangles = (uint32_t)s->job.scene.lightTris.size();
  info->has_environment = s->job.scene.hasEnvTexture ? 1 : 0;
  info->bvh_nodes = s->bvh().numNodes;
  info->bvh_depth = s->bvh().depth;
  info->device_bytes = s->sceneBytes + s->bvh().bytes;
  info->bvh_width = s->bvh().width;
  info->camera_exposure = s->job.scene.camera.exposure;
  return 0;
}

// ---- device films with snapshots for progressive display ----
struct HastyFilms {
  HastyScene* scene = nullptr;
  size_t floats = 0;           // W*H*3 per film
  float* film = nullptr;       // colour, normal, albedo (device)
  float* snap = nullptr;       // the last snapshot's copy (device)
  float* staging = nullptr;    // its copy in pinned host memory
  hipStream_t copyStream = nullptr;
  hipEvent_t rendered = nullptr;  // the last snapshot's device copy ended (scene stream)
  hipEvent_t copied = nullptr;    // the last snapshot's device-to-host copy ended
  float* estimate = nullptr;   // hasty_films_estimate's output (device, 3 films)
  uint32_t samples = 0;         // samples added to the films
  uint32_t snapSamples = 0;     // samples in the last snapshot
  bool unread = false;          // a snapshot not yet handed out
  std::mutex m;
  ~HastyFilms() {
    if (copyStream) (void)hipStreamSynchronize(copyStream);
    if (copied) (void)hipEventDestroy(copied);
    if (rendered) (void)hipEventDestroy(rendered);
    if (copyStream) (void)hipStreamDestroy(copyStream);
    if (staging) (void)hipHostFree(staging);
    if (snap) (void)hipFree(snap);
    if (film) (void)hipFree(film);
    if (estimate) (void)hipFree(estimate);
  }
};

int hasty_scene_set_settings(HastyScene* s, const HastyRenderSettings* st) {
  if (!s || !st) {
    g_lastError = "null argument";
    return 1;
  }
  hasty::RenderSettings r = s->job.settings;
  r.rouletteStartDepth = st->roulette_start_depth;
  r.rouletteQ = st->roulette_q;
  r.width = st->width;
  r.height = st->height;
  r.numThreads = st->num_threads;
  r.maxDepth = st->max_depth;
  r.numSamples = st->num_samples;
  r.useMIS = st->use_mis != 0;
  try {
    checkSettings(r);  // a rejected call leaves the scene unchanged
  } catch (const std::exception& e) {
    return fail(e, 1);
  }
  s->job.settings = r;
  s->applySettings();
  return 0;
}

int hasty_render_pt_device(HastyScene* s, uint32_t seed, uint32_t first, uint32_t n, float* dC, float* dN,
                           float* dA) {
  if (!s || !dC || !dN || !dA) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(s->device));
    if (n > 0) s->renderDevice(seed, first, n, dC, dN, dA);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_films_create(HastyScene* s, HastyFilms** out) {
  if (!s || !out) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(s->device));
    std::unique_ptr<HastyFilms> f(new HastyFilms());
    f->scene = s;
    f->floats = (size_t)s->dev.width * s->dev.height * 3;
    if (f->floats == 0) throw std::runtime_error("empty film");
    HIPCHECK(hipMalloc(&f->film, 3 * f->floats * sizeof(float)));
    HIPCHECK(hipMalloc(&f->snap, 3 * f->floats * sizeof(float)));
    HIPCHECK(hipHostMalloc(&f->staging, 3 * f->floats * sizeof(float), hipHostMallocDefault));
    HIPCHECK(hipStreamCreate(&f->copyStream));
    HIPCHECK(hipEventCreateWithFlags(&f->copied, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&f->rendered, hipEventDisableTiming));
    HIPCHECK(hipMemsetAsync(f->film, 0, 3 * f->floats * sizeof(float), s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    *out = f.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

void hasty_films_destroy(HastyFilms* f) { delete f; }

int hasty_films_render(HastyFilms* f, uint32_t seed, uint32_t first, uint32_t n) {
  if (!f) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HastyScene* s = f->scene;
    HIPCHECK(hipSetDevice(s->device));
    if ((size_t)s->dev.width * s->dev.height * 3 != f->floats)
      throw std::runtime_error("the scene's film size changed after the films were created");
    if (n > 0) s->renderDevice(seed, first, n, f->film, f->film + f->floats, f->film + 2 * f->floats);
    std::lock_guard<std::mutex> lock(f->m);
    f->samples += n;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_films_snapshot(HastyFilms* f) {
  if (!f) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HastyScene* s = f->scene;
    HIPCHECK(hipSetDevice(s->device));
    std::lock_guard<std::mutex> lock(f->m);
    const size_t bytes = 3 * f->floats * sizeof(float);
    // the previous snapshot's readback must have left `snap` before it is
    // overwritten; the copy follows the work queued on the scene's stream
    HIPCHECK(hipStreamWaitEvent(s->stream, f->copied, 0));
    HIPCHECK(hipMemcpyAsync(f->snap, f->film, bytes, hipMemcpyDeviceToDevice, s->stream));
    HIPCHECK(hipEventRecord(f->rendered, s->stream));
    HIPCHECK(hipStreamWaitEvent(f->copyStream, f->rendered, 0));
    HIPCHECK(hipMemcpyAsync(f->staging, f->snap, bytes, hipMemcpyDeviceToHost, f->copyStream));
    HIPCHECK(hipEventRecord(f->copied, f->copyStream));
    f->snapSamples = f->samples;
    f->unread = true;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_films_read(HastyFilms* f, int wait, float* color, float* normal, float* albedo, uint32_t* numSamples) {
  if (!f || !color || !normal || !albedo || !numSamples) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(f->scene->device));
    std::lock_guard<std::mutex> lock(f->m);
    if (!f->unread) return 4;
    if (!wait) {
      hipError_t q = hipEventQuery(f->copied);
      if (q == hipErrorNotReady) return 4;
      HIPCHECK(q);
    }
    HIPCHECK(hipEventSynchronize(f->copied));
    memcpy(color, f->staging, f->floats * sizeof(float));
    memcpy(normal, f->staging + f->floats, f->floats * sizeof(float));
    memcpy(albedo, f->staging + 2 * f->floats, f->floats * sizeof(float));
    *numSamples = f->snapSamples;
    f->unread = false;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_films_estimate(HastyFilms* const* films, uint32_t n, float exposure, float* color, float* normal,
                         float* albedo) {
  if (!films || n == 0 || !color || !normal || !albedo) {
    g_lastError = "null argument";
    return 1;
  }
  if (n > 64) {
    g_lastError = "1 to 64 film sets per estimate";
    return 1;
  }
  try {
    HastyFilms* f0 = films[0];
    if (!f0) throw std::runtime_error("null film set");
    HastyScene* s = f0->scene;
    HIPCHECK(hipSetDevice(s->device));
    std::vector<const float*> planes[3];
    std::vector<uint32_t> counts;
    for (uint32_t k = 0; k < n; ++k) {
      HastyFilms* f = films[k];
      if (!f || f->scene != s) throw std::runtime_error("film sets of one scene");
      std::lock_guard<std::mutex> lock(f->m);
      for (int p = 0; p < 3; ++p) planes[p].push_back(f->film + p * f->floats);
      counts.push_back(f->samples);
    }
    std::lock_guard<std::mutex> lock(f0->m);
    const size_t floats = f0->floats;
    if (!f0->estimate) HIPCHECK(hipMalloc(&f0->estimate, 3 * floats * sizeof(float)));
    // PathTracingViewer.cpp:159-162: the colour estimate times 2^exposure
    const float scale = exposure == 0.0f ? 1.0f : std::pow(2.0f, exposure);
    for (int p = 0; p < 3; ++p)
      hasty::launchEstimate(planes[p].data(), counts.data(), (int)n, p == 0 ? scale : 1.0f, floats,
                            f0->estimate + p * floats, s->stream);
    float* out[3] = {color, normal, albedo};
    for (int p = 0; p < 3; ++p)
      HIPCHECK(hipMemcpyAsync(out[p], f0->estimate + p * floats, floats * sizeof(float), hipMemcpyDeviceToHost,
                              s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_films_device(HastyFilms* f, float** color, float** normal, float** albedo, uint32_t* numSamples) {
  if (!f || !color || !normal || !albedo || !numSamples) {
    g_lastError = "null argument";
    return 1;
  }
  std::lock_guard<std::mutex> lock(f->m);
  *color = f->film;
  *normal = f->film + f->floats;
  *albedo = f->film + 2 * f->floats;
  *numSamples = f->samples;
  return 0;
}

int hasty_render_pt(HastyScene* s, uint32_t seed, uint32_t first, uint32_t n, float* color, float* normal,
                    float* albedo) {
  if (!s || !color || !normal || !albedo) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(s->device));
    size_t P = (size_t)s->dev.width * s->dev.height * 3;
    if (s->filmFloats < 3 * P) {
      if (s->dFilm) (void)hipFree(s->dFilm);
      HIPCHECK(hipMalloc(&s->dFilm, 3 * P * sizeof(float)));
      s->filmFloats = 3 * P;
    }
    float* dC = s->dFilm;
    float* dN = dC + P;
    float* dA = dN + P;
    HIPCHECK(hipMemcpy(dC, color, P * sizeof(float), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dN, normal, P * sizeof(float), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dA, albedo, P * sizeof(float), hipMemcpyHostToDevice));
    if (n > 0) s->renderDevice(seed, first, n, dC, dN, dA);
    HIPCHECK(hipMemcpy(color, dC, P * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(normal, dN, P * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(albedo, dA, P * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_trace_closest(HastyScene* s, const float* rays, uint32_t n, uint32_t* geomPrim, int32_t* tri, float* tuv) {
  if (!s || (n > 0 && (!rays || !geomPrim || !tri || !tuv))) {
    g_lastError = "null argument";
    return 1;
  }
  if (n == 0) return 0;
  try {
    HIPCHECK(hipSetDevice(s->device));
    ensureTraceOverflow(s);
    // rays, triangle ids and (t, u, v) share the scene's trace buffer
    char* io = (char*)s->traceIO.ensure((sizeof(float) * 6 + sizeof(int32_t) + sizeof(float) * 3) * (size_t)n);
    float* dRays = (float*)io;
    int32_t* dTri = (int32_t*)(io + sizeof(float) * 6 * (size_t)n);
    float* dTuv = (float*)(io + (sizeof(float) * 6 + sizeof(int32_t)) * (size_t)n);
    HIPCHECK(hipMemcpy(dRays, rays, sizeof(float) * 6 * n, hipMemcpyHostToDevice));
    hasty::launchTraceClosest(s->dev, dRays, n, dTri, dTuv, s->traceOverflow, s->traceOverflowThreads, s->stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s->stream));
    HIPCHECK(hipMemcpy(tri, dTri, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(tuv, dTuv, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) {
      if (tri[i] < 0) {
        geomPrim[2 * i] = geomPrim[2 * i + 1] = 0xFFFFFFFFu;
      } else {
        geomPrim[2 * i] = s->job.scene.triGeomID[tri[i]];
        geomPrim[2 * i + 1] = s->job.scene.triPrimID[tri[i]];
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

static_assert(sizeof(HastyHit) == sizeof(hasty::HitRecord), "HastyHit layout");
static_assert(sizeof(HastyRay) == 6 * sizeof(float), "HastyRay layout");

int hasty_trace_hits_device(HastyScene* s, const HastyRay* dRays, uint32_t n, HastyHit* dHits) {
  if (!s || (n > 0 && (!dRays || !dHits))) {
    g_lastError = "null argument";
    return 1;
  }
  if (((uintptr_t)dHits & 15u) != 0 || ((uintptr_t)dRays & 3u) != 0) {
    g_lastError = "hits must be 16-byte aligned and rays 4-byte aligned";
    return 1;
  }
  if (n == 0) return 0;
  try {
    HIPCHECK(hipSetDevice(s->device));
    ensureTraceOverflow(s);
    hasty::launchTraceHits(s->dev, (const float*)dRays, n, (hasty::HitRecord*)dHits, s->traceOverflow,
                           s->traceOverflowThreads, s->stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s->stream));
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_trace_hits(HastyScene* s, const HastyRay* rays, uint32_t n, HastyHit* hits) {
  if (!s || (n > 0 && (!rays || !hits))) {
    g_lastError = "null argument";
    return 1;
  }
  if (n == 0) return 0;
  try {
    HIPCHECK(hipSetDevice(s->device));
    // rays then hits in the scene's trace buffer (hits 256-byte aligned)
    const size_t rayBytes = (sizeof(HastyRay) * (size_t)n + 255) & ~(size_t)255;
    char* io = (char*)s->traceIO.ensure(rayBytes + sizeof(HastyHit) * (size_t)n);
    HIPCHECK(hipMemcpy(io, rays, sizeof(HastyRay) * (size_t)n, hipMemcpyHostToDevice));
    int rc = hasty_trace_hits_device(s, (const HastyRay*)io, n, (HastyHit*)(io + rayBytes));
    if (rc != 0) return rc;
    HIPCHECK(hipMemcpy(hits, io + rayBytes, sizeof(HastyHit) * (size_t)n, hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_set_profiling(HastyScene* s, int flags) {
  if (!s) return 1;
  s->profiling = (flags & 1) != 0;
  s->counting = (flags & 2) != 0;
  if (s->profiling) {
    if (!s->timer) s->timer = std::make_unique<EventTimer>();
    s->timer->only = (flags & 4) != 0 ? hasty::kKExtend : -1;
  }
  return 0;
}

int hasty_kernel_stats(HastyScene* s, HastyKernelStat* out, int maxStats) {
  if (!s || !out) return -1;
  if (!s->timer) return 0;
  try {
    s->timer->collect();
  } catch (const std::exception& e) {
    g_lastError = e.what();
    return -1;
  }
  int n = 0;
  for (int k = 0; k < hasty::kNumKernels && n < maxStats; ++k) {
    if (s->timer->launches[k] == 0) continue;
    memset(out[n].name, 0, sizeof(out[n].name));
    strncpy(out[n].name, hasty::kernelName(k), sizeof(out[n].name) - 1);
    out[n].launches = s->timer->launches[k];
    out[n].total_ms = s->timer->ms[k];
    ++n;
  }
  return n;
}

int hasty_reset_stats(HastyScene* s) {
  if (!s) return 1;
  if (s->timer) s->timer->reset();
  for (int i = 0; i < hasty::kNumStats; ++i) s->stats[i] = 0;
  return 0;
}

int hasty_counters_ex(HastyScene* s, uint64_t* out, uint32_t n) {
  if (!s || (n > 0 && !out)) return 1;
  const uint64_t v[7] = {s->stats[hasty::kStatExtRays],
                         s->stats[hasty::kStatShadowRays],
                         s->stats[hasty::kStatNodes] + s->stats[hasty::kStatShadowNodes],
                         s->stats[hasty::kStatTris] + s->stats[hasty::kStatShadowTris],
                         s->stats[hasty::kStatMediaOverflow],
                         s->stats[hasty::kStatShadowNodes],
                         s->stats[hasty::kStatShadowTris]};
  for (uint32_t i = 0; i < n && i < 7; ++i) out[i] = v[i];
  return 0;
}

int hasty_counters(HastyScene* s, uint64_t* out) { return hasty_counters_ex(s, out, 5); }

int hasty_set_launch(HastyScene* s, uint32_t slots, uint32_t traceBlocks, uint32_t shadeBlocks) {
  if (!s) return 1;
  if (slots) s->slotsWanted = slots;
  if (traceBlocks) s->launch.traceBlocks = traceBlocks;
  if (shadeBlocks) s->launch.shadeBlocks = shadeBlocks;
  return 0;
}

int hasty_get_launch(const HastyScene* s, HastyLaunch* out) {
  if (!s || !out) {
    g_lastError = "null argument";
    return 1;
  }
  out->path_slots = s->slotsWanted ? s->slotsWanted : (1u << 23);
  out->trace_blocks = s->launch.traceBlocks;
  out->shade_blocks = s->launch.shadeBlocks;
  out->chunk = s->launch.chunk;
  out->use_graph = s->launch.useGraph ? 1 : 0;
  out->bvh_width = (uint32_t)s->bvh().width;
  return 0;
}

int hasty_set_launch_ex(HastyScene* s, const HastyLaunch* l) {
  if (!s || !l) {
    g_lastError = "null argument";
    return 1;
  }
  if (l->path_slots) s->slotsWanted = l->path_slots;
  if (l->trace_blocks) s->launch.traceBlocks = l->trace_blocks;
  if (l->shade_blocks) s->launch.shadeBlocks = l->shade_blocks;
  if (l->chunk) s->launch.chunk = l->chunk;
  s->launch.useGraph = l->use_graph != 0;
  if (l->bvh_width) {
    if (l->bvh_width != 4 && l->bvh_width != 8) {
      g_lastError = "bvh_width must be 4 or 8";
      return 1;
    }
    try {
      HIPCHECK(hipSetDevice(s->device));
      s->useBvhWidth((int)l->bvh_width);
    } catch (const std::exception& e) {
      return fail(e, 3);
    }
  }
  return 0;
}

static int batchItems(const HastyBatchJob* jobs, uint32_t n, std::vector<BatchItem>& items) {
  if (!jobs || n == 0) {
    g_lastError = "null argument";
    return 1;
  }
  items.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    const HastyBatchJob& j = jobs[i];
    if (!j.scene || !j.d_color || !j.d_normal || !j.d_albedo) {
      g_lastError = "null argument";
      return 1;
    }
    items[i] = BatchItem{j.scene, j.seed, j.first_sample, j.num_samples, j.d_color, j.d_normal, j.d_albedo};
  }
  return 0;
}

int hasty_render_batch_device(const HastyBatchJob* jobs, uint32_t n) {
  std::vector<BatchItem> items;
  if (int rc = batchItems(jobs, n, items)) return rc;
  try {
    HIPCHECK(hipSetDevice(jobs[0].scene->device));
    renderBatch(jobs[0].scene, items.data(), n);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

// Autotuner: one pass of coordinate descent over the launch knobs, each
// candidate timed on a short render of the real jobs (trial_samples samples
// of each, into scratch films).  Images never depend on the knobs.
static int autotuneItems(std::vector<BatchItem> items, uint32_t trialSamples, HastyLaunch* chosen) {
  HastyScene* owner = items[0].scene;
  std::vector<void*> films;
  try {
    HIPCHECK(hipSetDevice(owner->device));
    for (BatchItem& it : items) {
      size_t P = (size_t)it.scene->dev.width * it.scene->dev.height * 3;
      void* d = nullptr;
      HIPCHECK(hipMalloc(&d, 3 * P * sizeof(float)));
      films.push_back(d);
      it.color = (float*)d;
      it.normal = it.color + P;
      it.albedo = it.normal + P;
      it.first = 0;
      it.n = std::max<uint32_t>(1, trialSamples);
    }
    bool prof = owner->profiling, cnt = owner->counting;
    owner->profiling = owner->counting = false;
    auto timeIt = [&]() {
      auto t = std::chrono::steady_clock::now();
      renderBatch(owner, items.data(), (uint32_t)items.size());
      return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    };
    HastyLaunch best;
    hasty_get_launch(owner, &best);
    hasty_set_launch_ex(owner, &best);
    timeIt();  // warm-up (allocations)
    double bestT = timeIt();
    const std::vector<uint32_t> slots = {1u << 20, 1u << 21, 1u << 22, 1u << 23};
    const std::vector<uint32_t> trace = {1024, 2048, 4096};
    const std::vector<uint32_t> shade = {2048, 4096, 8192};
    const std::vector<uint32_t> chunks = {4, 8, 16};
    auto tryKnob = [&](auto setter, const std::vector<uint32_t>& values) {
      HastyLaunch base = best;
      for (uint32_t v : values) {
        HastyLaunch c = base;
        setter(c, v);
        hasty_set_launch_ex(owner, &c);
        double t = timeIt();
        if (t < bestT) {
          bestT = t;
          best = c;
        }
      }
      hasty_set_launch_ex(owner, &best);
    };
    tryKnob([](HastyLaunch& c, uint32_t v) { c.path_slots = v; }, slots);
    tryKnob([](HastyLaunch& c, uint32_t v) { c.trace_blocks = v; }, trace);
    tryKnob([](HastyLaunch& c, uint32_t v) { c.shade_blocks = v; }, shade);
    tryKnob([](HastyLaunch& c, uint32_t v) { c.chunk = v; }, chunks);
    tryKnob([](HastyLaunch& c, uint32_t v) { c.use_graph = (int32_t)v; }, {0u, 1u});
    // node width: every job of the trial renders with the candidate width
    // (a width whose tree is too deep for the traversal stack is skipped)
    {
      HastyLaunch base = best;
      for (uint32_t wdt : {4u, 8u}) {
        bool ok = true;
        for (BatchItem& it : items) {
          try {
            it.scene->useBvhWidth((int)wdt);
          } catch (const std::exception&) {
            ok = false;
          }
        }
        if (!ok) continue;
        double t = timeIt();
        if (t < bestT) {
          bestT = t;
          best = base;
          best.bvh_width = wdt;
        }
      }
      const int chosenWidth = best.bvh_width ? (int)best.bvh_width : 4;
      for (BatchItem& it : items) it.scene->useBvhWidth(chosenWidth);
    }
    owner->profiling = prof;
    owner->counting = cnt;
    hasty_get_launch(owner, &best);
    if (chosen) *chosen = best;
    for (void* d : films) (void)hipFree(d);
    return 0;
  } catch (const std::exception& e) {
    for (void* d : films) (void)hipFree(d);
    return fail(e, 3);
  }
}

int hasty_autotune_batch(const HastyBatchJob* jobs, uint32_t n, uint32_t trialSamples, HastyLaunch* chosen) {
  if (!jobs || n == 0) {
    g_lastError = "null argument";
    return 1;
  }
  std::vector<BatchItem> items(n);
  for (uint32_t i = 0; i < n; ++i) {
    if (!jobs[i].scene) {
      g_lastError = "null argument";
      return 1;
    }
    items[i] = BatchItem{jobs[i].scene, jobs[i].seed, 0, trialSamples, nullptr, nullptr, nullptr};
  }
  return autotuneItems(items, trialSamples, chosen);
}

int hasty_autotune(HastyScene* s, uint32_t seed, uint32_t trialSamples, HastyLaunch* chosen) {
  if (!s) {
    g_lastError = "null argument";
    return 1;
  }
  return autotuneItems({BatchItem{s, seed, 0, trialSamples, nullptr, nullptr, nullptr}}, trialSamples, chosen);
}

int hasty_sppm_create(HastyScene* s, uint32_t seed, const HastySppmParams* params, HastySppm** out) {
  if (!s || !out) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(s->device));
    if (s->dev.width == 0 || s->dev.height == 0) throw std::runtime_error("empty film");
    // The photon mapper's visible points and photon records carry no texture
    // coordinates; textured scenes render with the path tracer only.
    for (const hasty::Material& m : s->job.scene.materials)
      if (m.textured()) throw std::runtime_error("photon mapping does not support texture maps (material '" + m.name + "')");
    hasty::SppmParams p;
    if (params) {
      p.alpha = params->alpha;
      p.footprintScale = params->footprint_scale;
      p.initialRadius = params->initial_radius;
      p.storeFraction = params->store_fraction;
      p.emitFraction = params->emit_fraction;
      p.fixedPointFlux = params->fixed_point_flux != 0;
      if (params->flux_frac_bits) p.fluxFracBits = params->flux_frac_bits;
      p.gatherShareCandidates = params->gather_share_candidates;
      p.gatherUnionFactor = params->gather_union_factor;
    }
    if (p.fluxFracBits < 0 || p.fluxFracBits > 40) throw std::runtime_error("flux_frac_bits must be in [0, 40]");
    if (!(p.alpha > 0.0f && p.alpha <= 1.0f)) throw std::runtime_error("alpha must be in (0, 1]");
    hdev::DevScene d = s->dev;
    d.seed = seed;
    auto h = std::make_unique<HastySppm>();
    h->scene = s;
    h->r = std::make_unique<hasty::SppmRenderer>(d, p, s->stream);
    *out = h.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 2);
  }
}

void hasty_sppm_destroy(HastySppm* h) { delete h; }

int hasty_sppm_pass_device(HastySppm* h, uint32_t pass, float* dC, float* dN, float* dA) {
  if (!h || !dC || !dN || !dA) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HastyScene* s = h->scene;
    HIPCHECK(hipSetDevice(s->device));
    float ox, oy;
    haltonPoint(pass, &ox, &oy);
    if (s->profiling && !s->timer) s->timer = std::make_unique<EventTimer>();
    h->r->pass(pass, ox, oy, dC, dN, dA, s->profiling ? s->timer.get() : nullptr);
    if (s->profiling) s->timer->collect();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_sppm_pass(HastySppm* h, uint32_t pass, float* color, float* normal, float* albedo) {
  if (!h || !color || !normal || !albedo) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HastyScene* s = h->scene;
    HIPCHECK(hipSetDevice(s->device));
    size_t P = (size_t)s->dev.width * s->dev.height * 3;
    if (!h->dFilm) HIPCHECK(hipMalloc(&h->dFilm, 3 * P * sizeof(float)));
    int rc = hasty_sppm_pass_device(h, pass, h->dFilm, h->dFilm + P, h->dFilm + 2 * P);
    if (rc != 0) return rc;
    HIPCHECK(hipMemcpy(color, h->dFilm, P * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(normal, h->dFilm + P, P * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(albedo, h->dFilm + 2 * P, P * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_sppm_info(const HastySppm* h, HastySppmInfo* info) {
  if (!h || !info) {
    g_lastError = "null argument";
    return 1;
  }
  const hasty::SppmPassInfo& i = h->r->lastInfo();
  info->emitted = i.emitted;
  info->stored = i.stored;
  info->cell_size = i.cellSize;
  info->passes_done = h->r->passesDone();
  info->emit_limit = h->r->emitLimit();
  info->store_limit = h->r->storeLimit();
  info->flux_overflows = i.fluxOverflows;
  info->gather_queries = i.gatherQueries;
  info->gather_candidates = i.gatherCandidates;
  info->gather_found = i.gatherFound;
  return 0;
}

int hasty_sppm_pixels(HastySppm* h, float* out) {
  if (!h || !out) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(h->scene->device));
    h->r->pixelStats(out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_sppm_radii(HastySppm* h, float* out) {
  if (!h || !out) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(h->scene->device));
    h->r->radii(out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_sppm_photons(HastySppm* h, float* out, uint64_t maxPhotons) {
  if (!h || !out) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(h->scene->device));
    h->r->photons(out, (size_t)maxPhotons);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

}  // extern "C"
// ------------------------------------------------------------ path guide ----
extern "C" {

uint32_t hasty_guide_num_params(void) { return (uint32_t)guide::kParams; }

int hasty_guide_create(HastyScene* s, uint32_t seed, HastyGuide** out) {
  if (!s || !out) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(s->device));
    auto h = std::make_unique<HastyGuide>();
    h->scene = s;
    h->g = std::make_unique<guide::Guide>(s->stream);
    std::vector<float> p(guide::kParams);
    guide::Guide::initParams(seed, p.data());
    h->g->setParams(p.data());
    // scene box from the triangle positions
    const std::vector<float>& tp = s->job.scene.triPos;
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (size_t i = 0; i + 2 < tp.size(); i += 3)
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], tp[i + c]);
        hi[c] = std::max(hi[c], tp[i + c]);
      }
    if (tp.empty()) lo[0] = lo[1] = lo[2] = 0.0f, hi[0] = hi[1] = hi[2] = 1.0f;
    h->g->setBox(lo, hi);
    *out = h.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 2);
  }
}

void hasty_guide_destroy(HastyGuide* g) { delete g; }

int hasty_guide_set_box(HastyGuide* g, const float* lo, const float* hi) {
  if (!g || !lo || !hi) {
    g_lastError = "null argument";
    return 1;
  }
  g->g->setBox(lo, hi);
  return 0;
}

int hasty_guide_set_params(HastyGuide* g, const float* p) {
  if (!g || !p) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    g->g->setParams(p);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_guide_get_params(HastyGuide* g, float* p) {
  if (!g || !p) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    g->g->getParams(p);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

// queries: n x 9 floats (position, normal, wo); raw: n x 48 floats.
int hasty_guide_infer(HastyGuide* g, const float* queries, uint32_t n, float* raw) {
  if (!g || (n && (!queries || !raw))) {
    g_lastError = "null argument";
    return 1;
  }
  if (n == 0) return 0;
  try {
    HIPCHECK(hipSetDevice(g->scene->device));
    std::vector<float4> in((size_t)n * 3);
    for (uint32_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k)
        in[3 * i + k] = make_float4(queries[9 * i + 3 * k], queries[9 * i + 3 * k + 1], queries[9 * i + 3 * k + 2], 0.0f);
    float4* d = g->buffer((size_t)n * 12 + (size_t)n * 48);
    float4* dRaw = d + (size_t)n * 3;
    hipStream_t st = g->g->stream();
    HIPCHECK(hipMemcpyAsync(d, in.data(), sizeof(float4) * in.size(), hipMemcpyHostToDevice, st));
    g->g->inferDevice(d, n, dRaw);
    HIPCHECK(hipMemcpyAsync(raw, dRaw, sizeof(float) * 48 * (size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

// samples: n x 13 floats (position, normal, wo, wi, weight F).  Adds the
// gradient of the mean loss -F log q(wi) to the guide's gradient buffer (the
// buffer is cleared first when `clear`), optionally copies it to grad_out,
// and returns the mean loss in *loss.
// Samples of `stride` floats (13: position, normal, wo, wi, F; 15: then the
// BSDF density of wi and the density wi was drawn with).
static void guideGradient(HastyGuide* g, const float* smp, uint32_t n, bool clear, float* gradOut, float* loss,
                          uint32_t stride = 13) {
  std::vector<float4> in((size_t)n * 4);
  for (uint32_t i = 0; i < n; ++i) {
    const float* s = smp + stride * (size_t)i;
    in[4 * i] = make_float4(s[0], s[1], s[2], s[12]);
    in[4 * i + 1] = make_float4(s[3], s[4], s[5], 0.0f);
    in[4 * i + 2] = make_float4(s[6], s[7], s[8], stride > 13 ? s[14] : 0.0f);
    in[4 * i + 3] = make_float4(s[9], s[10], s[11], stride > 13 ? s[13] : 0.0f);
  }
  float4* d = g->buffer((size_t)n * 16);
  hipStream_t st = g->g->stream();
  HIPCHECK(hipMemcpyAsync(d, in.data(), sizeof(float4) * in.size(), hipMemcpyHostToDevice, st));
  if (clear) g->g->zeroGrad();
  float l = g->g->gradientDevice(d, n);
  if (loss) *loss = l;
  if (gradOut) g->g->getGrad(gradOut);
}

int hasty_guide_gradient(HastyGuide* g, const float* samples, uint32_t n, float* grad, float* loss) {
  if (!g || !samples || n == 0) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(g->scene->device));
    guideGradient(g, samples, n, true, grad, loss);
    g->g->zeroGrad();
    HIPCHECK(hipStreamSynchronize(g->g->stream()));
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_guide_gradient_ex(HastyGuide* g, const float* samples, uint32_t n, uint32_t stride, float* grad,
                            float* loss) {
  if (!g || !samples || n == 0) {
    g_lastError = "null argument";
    return 1;
  }
  if (stride != 13 && stride != 15) {
    g_lastError = "sample stride must be 13 or 15 floats";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(g->scene->device));
    guideGradient(g, samples, n, true, grad, loss, stride);
    g->g->zeroGrad();
    HIPCHECK(hipStreamSynchronize(g->g->stream()));
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

int hasty_guide_set_selection(HastyGuide* g, float lo, float hi) {
  if (!g) {
    g_lastError = "null argument";
    return 1;
  }
  const bool off = lo == 0.0f && hi == 0.0f;
  if (!off && !(lo > 0.0f && lo < hi && hi < 1.0f)) {
    g_lastError = "selection range must satisfy 0 < lo < hi < 1 (or lo = hi = 0 for the fixed fraction)";
    return 1;
  }
  g->g->setSelection(lo, hi);
  return 0;
}

int hasty_scene_set_guide(HastyScene* s, HastyGuide* g, float fraction, float trainLr) {
  if (!s) {
    g_lastError = "null argument";
    return 1;
  }
  if (g && g->scene != s) {
    g_lastError = "the guide belongs to another scene";
    return 1;
  }
  if (g && !(fraction > 0.0f && fraction < 1.0f)) {
    g_lastError = "guide fraction must be in (0, 1)";
    return 1;
  }
  s->guide = g;
  s->guideFraction = fraction;
  s->guideLr = trainLr;
  return 0;
}

int hasty_scene_set_guide_training(HastyScene* s, uint32_t minibatch, float weightCap) {
  if (!s) {
    g_lastError = "null argument";
    return 1;
  }
  if (minibatch < 16 || minibatch > HastyScene::kGuideSampleCap || !(weightCap > 0.0f)) {
    g_lastError = "minibatch must be in [16, 2^21] and the weight cap positive";
    return 1;
  }
  s->guideMinibatch = minibatch;
  s->guideWeightCap = weightCap;
  return 0;
}

int hasty_scene_set_guide_depth(HastyScene* s, uint32_t maxDepth, float maxThroughput) {
  if (!s) {
    g_lastError = "null argument";
    return 1;
  }
  if (!(maxThroughput > 0.0f)) {
    g_lastError = "the throughput limit must be positive";
    return 1;
  }
  s->guideMaxDepth = maxDepth;
  s->guideMaxThroughput = maxThroughput;
  return 0;
}

int hasty_scene_set_guide_photons(HastyScene* s, uint32_t photons, float share) {
  if (!s) {
    g_lastError = "null argument";
    return 1;
  }
  if (photons > (1u << 26) || !(share > 0.0f && share < 1e6f)) {
    g_lastError = "at most 2^26 light paths per training step and a positive share";
    return 1;
  }
  s->guidePhotons = photons;
  s->guidePhotonShare = share;
  return 0;
}

int hasty_scene_guide_stats(const HastyScene* s, float* lastLoss, uint32_t* lastSamples) {
  if (!s) {
    g_lastError = "null argument";
    return 1;
  }
  if (lastLoss) *lastLoss = s->lastGuideLoss;
  if (lastSamples) *lastSamples = s->lastGuideSamples;
  return 0;
}

int hasty_guide_train_step(HastyGuide* g, const float* samples, uint32_t n, float lr, float* loss) {
  if (!g || !samples || n == 0) {
    g_lastError = "null argument";
    return 1;
  }
  try {
    HIPCHECK(hipSetDevice(g->scene->device));
    guideGradient(g, samples, n, true, nullptr, loss);
    g->g->adamStep(lr);
    HIPCHECK(hipStreamSynchronize(g->g->stream()));
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 3);
  }
}

}  // extern "C"

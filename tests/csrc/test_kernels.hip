// Test-only HIP entry points: run the renderer's device primitives (the
// deterministic math, the RNG, the GGX/BSDF and sampling routines of
// pt_device.h) on the GPU over host-supplied inputs, so the GPU tests can
// compare them bit for bit with the CPU oracle.  Built into
// tests/libhasty_testkernels.so; the renderer does not link it.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../hastyrenderer_amd/csrc/guide.h"
#include "../../hastyrenderer_amd/csrc/pt_device.h"

using namespace hdev;

namespace {

template <class T>
T* toDevice(const T* h, size_t n) {
  T* d = nullptr;
  if (hipMalloc(&d, sizeof(T) * (n ? n : 1)) != hipSuccess) return nullptr;
  if (n) (void)hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice);
  return d;
}
template <class T>
T* devAlloc(size_t n) {
  T* d = nullptr;
  if (hipMalloc(&d, sizeof(T) * (n ? n : 1)) != hipSuccess) return nullptr;
  return d;
}

__global__ void kMath(int fn, const float* a, const float* b, uint32_t n, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x = a[i], y = b[i], r;
  switch (fn) {
    case 0: r = hm::sinf(x); break;
    case 1: r = hm::cosf(x); break;
    case 2: r = hm::tanf(x); break;
    case 3: r = hm::atan2f(x, y); break;
    case 4: r = hm::acosf(x); break;
    case 5: r = hm::powf_pos(x, y); break;
    case 6: r = hm::fmodf(x, y); break;
    case 7: r = hm::atanf(x); break;
    case 8: r = hm::fsqrt(x); break;
    case 9: r = x / y; break;
    default: r = (float)((double)x * (double)y + 1.0); break;
  }
  out[i] = r;
}

__global__ void kPhilox(const uint32_t* in, uint32_t n, uint32_t* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t* c = in + i * 6;
  hrng::U32x4 r = hrng::philox4x32_10(c[0], c[1], c[2], c[3], c[4], c[5]);
  for (int k = 0; k < 4; ++k) out[i * 4 + k] = r.v[k];
}

__global__ void kVndf(const float* in, uint32_t n, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* a = in + i * 7;
  V3 r = sampleVNDF(v3(a[4], a[5], a[6]), a[2], a[3], a[0], a[1]);
  out[i * 3] = r.x, out[i * 3 + 1] = r.y, out[i * 3 + 2] = r.z;
}

// Same draw sequence as oracle_sampler (stream (seed, 0, 0, 99)), one lane.
__global__ void kSampler(int kind, uint32_t seed, uint32_t count, V3 n, V3 wo, float param, float* out) {
  if (blockIdx.x * blockDim.x + threadIdx.x != 0) return;
  hrng::Stream rng = hrng::make_stream(seed, 0, 0, 99);
  for (uint32_t i = 0; i < count; ++i) {
    float* o = out + (size_t)i * 5;
    if (kind == 0) {
      float pdf;
      V3 d = sampleCosHemisphere(rng, n, &pdf);
      o[0] = d.x, o[1] = d.y, o[2] = d.z, o[3] = pdf, o[4] = cosHemispherePdf(n, d);
    } else if (kind == 1) {
      V3 t = anyOrthonormal(n);
      V3 b = cross(n, t);
      float pdf;
      V3 d = sampleVNDFGlobal(rng, n, param, param, t, b, wo, &pdf);
      o[0] = d.x, o[1] = d.y, o[2] = d.z, o[3] = pdf, o[4] = vndfGlobalPdf(n, param, param, t, b, wo, d);
    } else if (kind == 3) {
      float x, y;
      const float half = ngonSector((int)param) * 0.5f;
      sampleNGon(rng, (int)param, 0.0f, hm::cosf(half), hm::sinf(half), &x, &y);
      float anglePerSector = (float)(2.0 * kPi / (int)param);
      float areaSector = 0.5 * hm::sinf(anglePerSector);
      float areaNGon = (int)param * areaSector;
      float pdf = 1.0 / areaNGon;
      o[0] = x, o[1] = y, o[2] = 0.0f, o[3] = pdf, o[4] = pdf;
    } else {
      float s = rng.uniform01f();
      float t = hm::fsqrt(rng.uniform01f());
      float b1 = 1.0f - t, b3 = t * s, b2 = fmaxf(0.0f, 1.0f - b1 - b3);
      o[0] = b1, o[1] = b2, o[2] = b3, o[3] = 0.0f, o[4] = 0.0f;
    }
  }
}

// Flat interaction with normal n at x = (0.3, 0.2, 0.1), as oracle_bsdf_eval.
__device__ Interaction flatInteraction(V3 n) {
  Interaction it;
  it.tri = 0;
  it.geomID = 0;
  it.mat = 0;
  it.h = 0.0f;
  it.x = v3(0.3f, 0.2f, 0.1f);
  it.ng = n;
  it.ns = n;
  it.tangent = orthonormalizedOrAny(n, v3(-it.x.z, 0.0f, it.x.x));
  it.bitangent = cross(n, it.tangent);
  return it;
}
__global__ void kBsdfEval(DevMaterial m, V3 n, const float* dirs, uint32_t count, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  Interaction it = flatInteraction(n);
  V3 wo = load3(dirs + i * 6), wi = load3(dirs + i * 6 + 3);
  BsdfEval e = bsdfEvaluate(m, it, wo, wi);
  float* o = out + i * 7;
  o[0] = e.fD.x, o[1] = e.fD.y, o[2] = e.fD.z, o[3] = e.fC.x, o[4] = e.fC.y, o[5] = e.fC.z;
  o[6] = bsdfPdf(m, it, wo, wi, 1.0f);
}
__global__ void kBsdfSample(DevMaterial m, V3 n, V3 wo, uint32_t seed, uint32_t count, float* out) {
  if (blockIdx.x * blockDim.x + threadIdx.x != 0) return;
  Interaction it = flatInteraction(n);
  hrng::Stream rng = hrng::make_stream(seed, 0, 0, 97);
  for (uint32_t i = 0; i < count; ++i) {
    BsdfSample s = bsdfSample(m, rng, it, wo, -1, 1.0f, false, 3);
    float* o = out + (size_t)i * 7;
    V3 t = s.tD + s.tC;
    o[0] = s.dir.x, o[1] = s.dir.y, o[2] = s.dir.z, o[3] = t.x, o[4] = t.y, o[5] = t.z, o[6] = s.pdf;
  }
}

DevMaterial makeMaterial(const float* p) {
  // albedo rgb, roughness, metallic, specular, ior, anisotropy, transmission
  DevMaterial m{};
  m.albedo[0] = p[0], m.albedo[1] = p[1], m.albedo[2] = p[2];
  m.roughness = p[3];
  float alpha = p[3] * p[3];
  if (alpha < 1e-2f && alpha != 0.0f) alpha = 1e-2f;
  m.alpha = alpha;
  m.metallic = p[4];
  m.specular = p[5];
  m.ior = p[6];
  m.anisotropy = p[7];
  m.transmission = p[8];
  m.varyIOR = 0;
  m.isLight = 0;
  m.hasDiffuseLobe = ((1.0f - 0.5f * m.specular) * (1.0 - m.transmission) * (1.0 - m.metallic) > 0.0f) ? 1 : 0;
  return m;
}

int finish(float* dOut, float* hOut, size_t n) {
  if (hipDeviceSynchronize() != hipSuccess) return 2;
  if (hipMemcpy(hOut, dOut, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess) return 3;
  return 0;
}

}  // namespace

// One v_mfma_f32_16x16x16_bf16 of A (16x16) and B (16x16), both given
// row-major as floats and rounded to BF16, with the operand lane layout the
// guide kernels assume (A: i = l%16, k = 4(l/16)+v; B: j = l%16,
// k = 4(l/16)+v; D: j = l%16, i = 4(l/16)+v).  One wave, in bounds.
typedef short tkbf16x4 __attribute__((ext_vector_type(4)));
typedef float tkf32x4 __attribute__((ext_vector_type(4)));
__device__ inline short tkBf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__global__ void kMfma(const float* A, const float* B, float* D) {
  int l = threadIdx.x, r = l & 15, q = l >> 4;
  tkbf16x4 a, b;
  a.x = tkBf16(A[r * 16 + 4 * q + 0]);
  a.y = tkBf16(A[r * 16 + 4 * q + 1]);
  a.z = tkBf16(A[r * 16 + 4 * q + 2]);
  a.w = tkBf16(A[r * 16 + 4 * q + 3]);
  b.x = tkBf16(B[(4 * q + 0) * 16 + r]);
  b.y = tkBf16(B[(4 * q + 1) * 16 + r]);
  b.z = tkBf16(B[(4 * q + 2) * 16 + r]);
  b.w = tkBf16(B[(4 * q + 3) * 16 + r]);
  tkf32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  D[(4 * q + 0) * 16 + r] = acc.x;
  D[(4 * q + 1) * 16 + r] = acc.y;
  D[(4 * q + 2) * 16 + r] = acc.z;
  D[(4 * q + 3) * 16 + r] = acc.w;
}

// Mixture tables of n guide queries (x, n, wo: 9 floats each) by the path
// tracer's inference code: the network staged in LDS in lobe-major order,
// guide::tileOutputs on 16-query tiles, guide::lobeEntry.  Per query 10
// float4 (lobes 0..7 as (kappa mu, log(w norm) - kappa), then the 8 weights)
// and the selection logit.  A wave takes tiles t = wave, wave + waves, ...
// (uniform over the wave); reads and writes only queries e < n.
__global__ void __launch_bounds__(256) kGuideTables(guide::Net net, const float* q9, uint32_t n, float4* tables,
                                                    float* logits) {
  __shared__ guide::LdsNet ldsNet;
  guide::stageNet(net, &ldsNet, true);
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t t = wave; t * 16u < n; t += waves) {
    const uint32_t e = t * 16u + (uint32_t)r;
    V3 x = v3(0.0f, 0.0f, 0.0f), nn = v3(0.0f, 0.0f, 1.0f), wo = v3(0.0f, 0.0f, 1.0f);
    if (e < n) {
      const float* p = q9 + (size_t)e * 9;
      x = guide::normPos(net, v3(p[0], p[1], p[2]));
      nn = v3(p[3], p[4], p[5]);
      wo = v3(p[6], p[7], p[8]);
    }
    float r10[10], w0, w1, logit;
    guide::tileOutputs(&ldsNet, x, nn, wo, r10, &w0, &w1, &logit);
    if (e < n) {
      float4* out = tables + (size_t)e * 10;
      out[2 * q] = guide::lobeEntry(r10, w0);
      out[2 * q + 1] = guide::lobeEntry(r10 + 5, w1);
      reinterpret_cast<float2*>(out + guide::kLobes)[q] = make_float2(w0, w1);
      if (q == 0) logits[e] = logit;
    }
  }
}

// Per query i (its table at tables[10 i ..], as kGuideTables writes it): m
// directions sampled by guide::mixtureSampleTable from xi[3 t ..] and
// guide::mixturePdfTable at the directions evalDirs[3 t ..], t = i m + j.
// One thread per (query, sample), in bounds.
__global__ void kGuideMixture(const float4* tables, uint32_t total, uint32_t m, const float* xi, float* dirs,
                              const float* evalDirs, float* pdfs) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const float4* g = tables + (size_t)(t / m) * 10;
  const float* x = xi + (size_t)t * 3;
  const V3 d = guide::mixtureSampleTable(g, x[0], x[1], x[2]);
  dirs[(size_t)t * 3] = d.x;
  dirs[(size_t)t * 3 + 1] = d.y;
  dirs[(size_t)t * 3 + 2] = d.z;
  const float* e = evalDirs + (size_t)t * 3;
  pdfs[t] = guide::mixturePdfTable(g, v3(e[0], e[1], e[2]));
}

extern "C" {

int tk_mfma(const float* A, const float* B, float* D) {
  float* dA = toDevice(A, 256);
  float* dB = toDevice(B, 256);
  float* dD = devAlloc<float>(256);
  if (!dA || !dB || !dD) return 1;
  kMfma<<<1, 64>>>(dA, dB, dD);
  int rc = finish(dD, D, 256);
  (void)hipFree(dA), (void)hipFree(dB), (void)hipFree(dD);
  return rc;
}

int tk_math(int fn, const float* a, const float* b, uint32_t n, float* out) {
  float* da = toDevice(a, n);
  float* db = toDevice(b, n);
  float* dout = devAlloc<float>(n);
  if (!da || !db || !dout) return 1;
  kMath<<<(n + 255) / 256, 256>>>(fn, da, db, n, dout);
  int rc = finish(dout, out, n);
  (void)hipFree(da), (void)hipFree(db), (void)hipFree(dout);
  return rc;
}

int tk_philox(const uint32_t* in6, uint32_t n, uint32_t* out4) {
  uint32_t* din = toDevice(in6, (size_t)n * 6);
  uint32_t* dout = devAlloc<uint32_t>((size_t)n * 4);
  if (!din || !dout) return 1;
  kPhilox<<<(n + 255) / 256, 256>>>(din, n, dout);
  if (hipDeviceSynchronize() != hipSuccess) return 2;
  (void)hipMemcpy(out4, dout, sizeof(uint32_t) * n * 4, hipMemcpyDeviceToHost);
  (void)hipFree(din), (void)hipFree(dout);
  return 0;
}

int tk_vndf(const float* in7, uint32_t n, float* out3) {
  float* din = toDevice(in7, (size_t)n * 7);
  float* dout = devAlloc<float>((size_t)n * 3);
  if (!din || !dout) return 1;
  kVndf<<<(n + 255) / 256, 256>>>(din, n, dout);
  int rc = finish(dout, out3, (size_t)n * 3);
  (void)hipFree(din), (void)hipFree(dout);
  return rc;
}

int tk_sampler(int kind, uint32_t seed, uint32_t count, const float* n, const float* wo, float param, float* out5) {
  float* dout = devAlloc<float>((size_t)count * 5);
  if (!dout) return 1;
  kSampler<<<1, 64>>>(kind, seed, count, v3(n[0], n[1], n[2]), v3(wo[0], wo[1], wo[2]), param, dout);
  int rc = finish(dout, out5, (size_t)count * 5);
  (void)hipFree(dout);
  return rc;
}

int tk_bsdf_eval(const float* matParams, const float* n, const float* dirs6, uint32_t count, float* out7) {
  float* dd = toDevice(dirs6, (size_t)count * 6);
  float* dout = devAlloc<float>((size_t)count * 7);
  if (!dd || !dout) return 1;
  kBsdfEval<<<(count + 255) / 256, 256>>>(makeMaterial(matParams), v3(n[0], n[1], n[2]), dd, count, dout);
  int rc = finish(dout, out7, (size_t)count * 7);
  (void)hipFree(dd), (void)hipFree(dout);
  return rc;
}

int tk_bsdf_sample(const float* matParams, const float* n, const float* wo, uint32_t seed, uint32_t count,
                   float* out7) {
  float* dout = devAlloc<float>((size_t)count * 7);
  if (!dout) return 1;
  kBsdfSample<<<1, 64>>>(makeMaterial(matParams), v3(n[0], n[1], n[2]), v3(wo[0], wo[1], wo[2]), seed, count, dout);
  int rc = finish(dout, out7, (size_t)count * 7);
  (void)hipFree(dout);
  return rc;
}

// wT: the BF16 W^T of every layer (guide::kWeights), params: the FP32
// parameters (guide::kParams; biases read from them), lo3 / inv3: the
// position normalisation (x - lo) * inv.
int tk_guide_tables(const uint16_t* wT, uint32_t nWeights, const float* params, uint32_t nParams, const float* lo3,
                    const float* inv3, const float* q9, uint32_t n, float* tables, float* logits) {
  if (nWeights != (uint32_t)guide::kWeights || nParams != (uint32_t)guide::kParams || n == 0) return 4;
  short* dW = toDevice(reinterpret_cast<const short*>(wT), (size_t)nWeights);
  float* dP = toDevice(params, (size_t)nParams);
  float* dQ = toDevice(q9, (size_t)n * 9);
  float4* dT = devAlloc<float4>((size_t)n * 10);
  float* dL = devAlloc<float>(n);
  int rc = 1;
  if (dW && dP && dQ && dT && dL) {
    guide::Net net;
    memset(&net, 0, sizeof(net));
    net.wT = dW;
    net.w = nullptr;
    net.param = dP;
    for (int i = 0; i < 3; ++i) net.boxLo[i] = lo3[i], net.boxInv[i] = inv3[i];
    const uint32_t blocks = std::min<uint32_t>((n + 63) / 64, 64u);
    kGuideTables<<<blocks, 256>>>(net, dQ, n, dT, dL);
    rc = finish(dL, logits, n);
    if (rc == 0 && hipMemcpy(tables, dT, sizeof(float4) * (size_t)n * 10, hipMemcpyDeviceToHost) != hipSuccess) rc = 3;
  }
  (void)hipFree(dW), (void)hipFree(dP), (void)hipFree(dQ), (void)hipFree(dT), (void)hipFree(dL);
  return rc;
}

// nq tables (10 float4 each), m samples per table: xi and evalDirs hold
// 3 nq m floats, dirs receives 3 nq m, pdfs nq m.
int tk_guide_mixture(const float* tables, uint32_t nq, uint32_t m, const float* xi, float* dirs,
                     const float* evalDirs, float* pdfs) {
  const uint64_t total64 = (uint64_t)nq * m;
  if (nq == 0 || m == 0 || total64 > (1u << 24)) return 4;
  const uint32_t total = (uint32_t)total64;
  float4* dT = toDevice(reinterpret_cast<const float4*>(tables), (size_t)nq * 10);
  float* dX = toDevice(xi, (size_t)total * 3);
  float* dE = toDevice(evalDirs, (size_t)total * 3);
  float* dD = devAlloc<float>((size_t)total * 3);
  float* dP = devAlloc<float>(total);
  int rc = 1;
  if (dT && dX && dE && dD && dP) {
    kGuideMixture<<<(total + 255) / 256, 256>>>(dT, total, m, dX, dD, dE, dP);
    rc = finish(dP, pdfs, total);
    if (rc == 0 && hipMemcpy(dirs, dD, sizeof(float) * (size_t)total * 3, hipMemcpyDeviceToHost) != hipSuccess) rc = 3;
  }
  (void)hipFree(dT), (void)hipFree(dX), (void)hipFree(dE), (void)hipFree(dD), (void)hipFree(dP);
  return rc;
}

}  // extern "C"

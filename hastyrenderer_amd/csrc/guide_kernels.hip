// MLP path guide on gfx950: inference, fused forward + loss + backward, weight
// gradients and Adam, all with BF16 MFMA (v_mfma_f32_16x16x16_bf16) and FP32
// accumulation.
//
// Every layer is computed transposed, H_out^T (out x 16 queries) =
// W^T (out x in) . H_in^T (in x 16), by 16x16x16 tiles.  MFMA operand lanes:
//   A (16 x 16, row i, column k): lane l holds i = l % 16, k = 4 (l / 16) + v
//   B (16 x 16, row k, column j): lane l holds j = l % 16, k = 4 (l / 16) + v
//   D (16 x 16, row i, column j): lane l holds j = l % 16, i = 4 (l / 16) + v
// so an output tile in D layout is, after bias/ReLU and rounding, already the
// B operand of the next layer for the matching k-step: activations never
// leave registers between layers.  The backward pass uses the same trick with
// W (in x out) as the A operand.
#include "guide_kernels.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>
#include <string>

namespace guide {
using namespace hdev;

namespace {

void hipOk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------- infer ----
// Raw outputs (48 floats) of n queries given as (pos, normal, wo) float4
// triples; queries are taken 16 per wave.
// (The compiler keeps the loop-invariant LDS weights in registers here.)
__global__ void __launch_bounds__(256) kInfer(Net net, const float4* in, uint32_t n, float4* raw) {
  __shared__ LdsNet ldsNet;
  stageNet(net, &ldsNet);
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t t = wave; t * 16 < n; t += waves) {
    uint32_t j = t * 16 + r;
    uint32_t jj = j < n ? j : n - 1;
    float4 p4 = in[3 * jj], n4 = in[3 * jj + 1], w4 = in[3 * jj + 2];
    bf16x4 h0[5];
    inputOperands(normPos(net, f4xyz(p4)), f4xyz(n4), f4xyz(w4), h0);
    f32x4 o[3];
    forwardLds(&ldsNet, h0, o);
    if (j < n) {
#pragma unroll
      for (int mt = 0; mt < 3; ++mt) raw[(size_t)j * 12 + 4 * mt + q] = make_float4(o[mt].x, o[mt].y, o[mt].z, o[mt].w);
    }
  }
}

// -------------------------------------------------------- loss gradient ----
// d loss / d raw of one sample (lane-local, 41 outputs), loss = -F log q
// and, for a sample that carries the BSDF's density pB of wi and the density
// pS it was drawn with (both > 0) while the selection is learned (selLo <
// selHi), also F^2 pS / (c q + (1 - c) pB) with q held fixed: the second
// moment of the vertex's one-sample estimate (F = f L / pS), whose gradient
// moves only the selection logit.  Unlike the KL divergence, it weighs the
// rare samples that only the BSDF finds by their squared contribution, so
// the selection falls where the guide misses part of the integrand.
__device__ inline float lossGrad(const Net& net, const float* raw, V3 d, float F, float pB, float pS, float scale,
                                 float* g) {
  Mixture m = mixtureFromRaw(raw);
  float v[kLobes];
  float qsum = 0.0f;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) {
    v[k] = vmfPdf(m.mu[k], m.kappa[k], d);
    qsum += m.w[k] * v[k];
  }
  qsum = fmaxf(qsum, 1e-30f);  // all lobes pointing away: keep log finite
  float loss = -F * logf(qsum);
#pragma unroll
  for (int k = 0; k < kLobes; ++k) {
    float s = m.w[k] * v[k] / qsum;
    float kap = m.kappa[k];
    float cosd = dot(m.mu[k], d);
    float dlogv = 1.0f / kap - 2.0f * __expf(-2.0f * kap) / -expm1f(-2.0f * kap) + (cosd - 1.0f);
    float c = raw[5 * k + 3];
    float gc = (c > kCLo && c < kCHi) ? s * dlogv * kap : 0.0f;
    V3 rr = v3(raw[5 * k], raw[5 * k + 1], raw[5 * k + 2]);
    float nr = fmaxf(sqrtf(dot(rr, rr)), 1e-12f);
    V3 proj = d - m.mu[k] * cosd;
    float f = s * kap / nr;
    float sc = -F * scale;
    g[5 * k + 0] = sc * f * proj.x;
    g[5 * k + 1] = sc * f * proj.y;
    g[5 * k + 2] = sc * f * proj.z;
    g[5 * k + 3] = sc * gc;
    g[5 * k + 4] = sc * (s - m.w[k]);
  }
  g[kSelOut] = 0.0f;
  if (net.selHi > net.selLo && pB > 0.0f && pS > 0.0f) {
    const float z = raw[kSelOut];
    const float sg = 1.0f / (1.0f + __expf(-z));
    const float c = net.selLo + (net.selHi - net.selLo) * sg;
    const float pm = fmaxf(c * qsum + (1.0f - c) * pB, 1e-30f);
    const float m2 = F * F * pS / pm;
    loss += m2;
    g[kSelOut] = -scale * m2 * (qsum - pB) / pm * (net.selHi - net.selLo) * sg * (1.0f - sg);
  }
  return loss;
}

// Backward of one layer: dIn (D layout, MI tiles of the input dimension) =
// W (in x out) . G (out x 16), with G given as KS B operands.
template <int KS, int MI>
__device__ inline void layerBwd(const short* w, int dout, const bf16x4* g, f32x4* dIn) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int mi = 0; mi < MI; ++mi) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const short* row = w + (size_t)(16 * mi + r) * dout + 4 * q;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) acc = mfma(load4(row + 16 * ks), g[ks], acc);
    dIn[mi] = acc;
  }
}

// Stores a layer's B operands (features x 16 samples) feature-major:
// dst[f * batch + b0 + j].
template <int KS>
__device__ inline void storeOperands(short* dst, uint32_t batch, uint32_t b0, const bf16x4* x, bool valid) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  if (!valid) return;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    size_t f = 16 * ks + 4 * q;
    dst[(f + 0) * batch + b0 + r] = x[ks].x;
    dst[(f + 1) * batch + b0 + r] = x[ks].y;
    dst[(f + 2) * batch + b0 + r] = x[ks].z;
    dst[(f + 3) * batch + b0 + r] = x[ks].w;
  }
}

__device__ inline bf16x4 maskPack(f32x4 g, f32x4 z) {
  return pack4(z.x > 0.0f ? g.x : 0.0f, z.y > 0.0f ? g.y : 0.0f, z.z > 0.0f ? g.z : 0.0f, z.w > 0.0f ? g.w : 0.0f);
}

// ----------------------------------------------------------- train pass ----
// Samples: 4 float4 each (pos + F, normal, wo + the density wi was drawn
// with or 0, wi + the BSDF density of wi or 0).  Writes every layer's
// input operands H_l and output gradients G_l (BF16, feature-major) for the
// weight-gradient GEMM, and accumulates the loss.  batch = padded count (a
// multiple of 16); samples >= n have zero weight.
__global__ void __launch_bounds__(256) kTrainFwdBwd(Net net, const float4* smp, uint32_t n, uint32_t batch,
                                                     float scale, short* H, short* G, float* lossOut) {
  __shared__ float ldsRaw[4][16][kOutPad + 1];
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4, wv = threadIdx.x >> 6;
  float lossAcc = 0.0f;
  // groups of 4 tiles (one per wave); the trip count is uniform over the
  // block so that its barriers line up
  const uint32_t tiles = batch / 16;
  const uint32_t wpb = blockDim.x >> 6;
  for (uint32_t grp = blockIdx.x; grp * wpb < tiles; grp += gridDim.x) {
    uint32_t t = grp * wpb + wv;
    bool tileValid = t < tiles;
    uint32_t b0 = t * 16;
    uint32_t j = b0 + r;
    bool valid = tileValid && j < n;
    uint32_t jj = valid ? j : 0u;
    float4 p4 = smp[4 * jj], n4 = smp[4 * jj + 1], w4 = smp[4 * jj + 2], d4 = smp[4 * jj + 3];
    float F = valid ? p4.w : 0.0f;
    Acts a;
    inputOperands(normPos(net, f4xyz(p4)), f4xyz(n4), f4xyz(w4), a.h[0]);
    f32x4 o[3];
    forwardAll(net, a, o);
    // layer inputs for dW
    if (tileValid) {
      storeOperands<5>(H + (size_t)hOffset(0) * batch, batch, b0, a.h[0], true);
#pragma unroll
      for (int l = 1; l < kLayers; ++l) storeOperands<4>(H + (size_t)hOffset(l) * batch, batch, b0, a.h[l], true);
    }
    // raw outputs -> per-sample loss gradient (lanes 0..15 own one sample)
#pragma unroll
    for (int mt = 0; mt < 3; ++mt) {
      ldsRaw[wv][r][16 * mt + 4 * q + 0] = o[mt].x;
      ldsRaw[wv][r][16 * mt + 4 * q + 1] = o[mt].y;
      ldsRaw[wv][r][16 * mt + 4 * q + 2] = o[mt].z;
      ldsRaw[wv][r][16 * mt + 4 * q + 3] = o[mt].w;
    }
    __syncthreads();
    if (q == 0) {
      float raw[kOutUsed], g[kOutUsed];
#pragma unroll
      for (int i = 0; i < kOutUsed; ++i) raw[i] = ldsRaw[wv][r][i];
      float l = lossGrad(net, raw, f4xyz(d4), F, d4.w, w4.w, scale, g);
      if (valid) lossAcc += l;
#pragma unroll
      for (int i = 0; i < kOutUsed; ++i) ldsRaw[wv][r][i] = valid ? g[i] : 0.0f;
#pragma unroll
      for (int i = kOutUsed; i < kOutPad; ++i) ldsRaw[wv][r][i] = 0.0f;
    }
    __syncthreads();
    bf16x4 g4[4];
#pragma unroll
    for (int mt = 0; mt < 3; ++mt) {
      const float* src = &ldsRaw[wv][r][16 * mt + 4 * q];
      g4[mt] = pack4(src[0], src[1], src[2], src[3]);
    }
    __syncthreads();
    // output layer gradient G_4, then back through the hidden layers; the
    // selection output (row 40: tile 2, lane group 2, first value) trains its
    // own weights only, not the shared hidden layers
    if (tileValid) storeOperands<3>(G + (size_t)gOffset(4) * batch, batch, b0, g4, true);
    if (q == 2) g4[2].x = 0;
    f32x4 dh[4];
    layerBwd<3, 4>(net.w + wqOffset(4), kOutPad, g4, dh);
#pragma unroll
    for (int l = kLayers - 2; l >= 0; --l) {
      bf16x4 gl[4];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) gl[mt] = maskPack(dh[mt], a.z[l][mt]);
      if (tileValid) storeOperands<4>(G + (size_t)gOffset(l) * batch, batch, b0, gl, true);
      if (l > 0) layerBwd<4, 4>(net.w + wqOffset(l), kHidden, gl, dh);
    }
  }
  // loss: lanes 0..15 hold partial sums
  for (int off = 32; off > 0; off >>= 1) lossAcc += __shfl_xor(lossAcc, off);
  if (lane == 0 && lossAcc != 0.0f) atomicAdd(lossOut, lossAcc);
}

// ------------------------------------------------------ weight gradients ----
// dW^T[l] (out x in) = sum_b G_l[out][b] H_l[in][b] and db[l][out] =
// sum_b G_l[out][b]: one wave per (layer, 16x16 tile, batch chunk), FP32
// atomics into grad.
struct TileJob {
  int layer, mo, mi;
};
__global__ void __launch_bounds__(256) kWeightGrad(const short* H, const short* G, uint32_t batch, uint32_t chunk,
                                                   const TileJob* jobs, uint32_t nJobs, float* grad) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  uint32_t chunks = batch / chunk;
  if (wave >= nJobs * chunks) return;
  TileJob tj = jobs[wave / chunks];
  uint32_t c0 = (wave % chunks) * chunk;
  const int din = layerIn(tj.layer), dout = layerOut(tj.layer);
  const short* g = G + (size_t)gOffset(tj.layer) * batch + (size_t)(16 * tj.mo + r) * batch + c0 + 4 * q;
  const short* h = H + (size_t)hOffset(tj.layer) * batch + (size_t)(16 * tj.mi + r) * batch + c0 + 4 * q;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;
  for (uint32_t k = 0; k < chunk; k += 16) {
    bf16x4 ga = load4(g + k), hb = load4(h + k);
    acc = mfma(ga, hb, acc);
    if (tj.mi == 0) bsum += fromBf16(ga.x) + fromBf16(ga.y) + fromBf16(ga.z) + fromBf16(ga.w);
  }
  float* gw = grad + wOffset(tj.layer);
  int col = 16 * tj.mi + r;
  atomicAdd(&gw[(size_t)(16 * tj.mo + 4 * q + 0) * din + col], acc.x);
  atomicAdd(&gw[(size_t)(16 * tj.mo + 4 * q + 1) * din + col], acc.y);
  atomicAdd(&gw[(size_t)(16 * tj.mo + 4 * q + 2) * din + col], acc.z);
  atomicAdd(&gw[(size_t)(16 * tj.mo + 4 * q + 3) * din + col], acc.w);
  if (tj.mi == 0) {
    // lanes l and l + 16k hold parts of the same output row 16 mo + r
    bsum += __shfl_xor(bsum, 16);
    bsum += __shfl_xor(bsum, 32);
    if (q == 0) atomicAdd(&grad[bOffset(tj.layer) + 16 * tj.mo + r], bsum);
  }
  (void)dout;
}

// ------------------------------------------------------------------ Adam ----
// One Adam step of every parameter (the gradient is consumed and cleared),
// with the BF16 W^T and W copies of the weights refreshed in the same pass.
__global__ void __launch_bounds__(256) kAdam(float* param, float* m, float* v, float* grad, float lr, float b1,
                                             float b2, float eps, float c1, float c2, short* wT, short* w) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)kParams; i += gridDim.x * blockDim.x) {
    float g = grad[i];
    float mi = b1 * m[i] + (1.0f - b1) * g;
    float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    float p = param[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    param[i] = p;
    grad[i] = 0.0f;
    // BF16 copies of a weight (not of a bias), as kQuantize writes them
    int l = 0;
    while (l + 1 < kLayers && (uint32_t)wOffset(l + 1) <= i) ++l;
    const int din = layerIn(l), dout = layerOut(l);
    const uint32_t local = i - (uint32_t)wOffset(l);
    if (local < (uint32_t)(din * dout)) {
      const int o = local / din, k = local - o * din;
      const short b = toBf16(p);
      wT[wqOffset(l) + local] = b;
      w[wqOffset(l) + (size_t)k * dout + o] = b;
    }
  }
}
// BF16 copies of the weights: W^T as stored, W transposed.
__global__ void __launch_bounds__(256) kQuantize(const float* param, short* wT, short* w) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)kWeights; i += gridDim.x * blockDim.x) {
    int l = 0;
    while (l + 1 < kLayers && (uint32_t)wqOffset(l + 1) <= i) ++l;
    int din = layerIn(l), dout = layerOut(l);
    uint32_t local = i - wqOffset(l);
    int o = local / din, k = local - o * din;
    short b = toBf16(param[wOffset(l) + local]);
    wT[i] = b;
    w[wqOffset(l) + (size_t)k * dout + o] = b;
  }
}
// Training-sample weights: their sum, then F <- min(F * inv, cap).
__global__ void __launch_bounds__(256) kWeightSum(const float4* smp, uint32_t n, float* sum) {
  float a = 0.0f;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a += smp[4 * i].w;
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(sum, a);
}
__global__ void __launch_bounds__(256) kWeightScale(float4* smp, uint32_t n, float inv, float cap) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    smp[4 * i].w = fminf(smp[4 * i].w * inv, cap);
}
// dst[i] = src[(a i + b) mod n] with a coprime to n: a permutation of the
// 4-float4 samples.
__global__ void __launch_bounds__(256) kPermute(const float4* src, float4* dst, uint32_t n, uint32_t a, uint32_t b) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t j = (uint32_t)(((uint64_t)a * i + b) % n);
#pragma unroll
    for (int k = 0; k < 4; ++k) dst[(size_t)i * 4 + k] = src[(size_t)j * 4 + k];
  }
}
__global__ void kZero(float* a, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0.0f;
}

}  // namespace

// ====================================================================== host
Guide::Guide(hipStream_t stream) : stream_(stream) {
  hipOk(hipMalloc(&dParam_, sizeof(float) * kParams), "hipMalloc");
  hipOk(hipMalloc(&dM_, sizeof(float) * kParams), "hipMalloc");
  hipOk(hipMalloc(&dV_, sizeof(float) * kParams), "hipMalloc");
  hipOk(hipMalloc(&dGrad_, sizeof(float) * kParams), "hipMalloc");
  hipOk(hipMalloc(&dWT_, sizeof(short) * kWeights), "hipMalloc");
  hipOk(hipMalloc(&dW_, sizeof(short) * kWeights), "hipMalloc");
  hipOk(hipMalloc(&dLoss_, sizeof(float)), "hipMalloc");
  std::vector<TileJob> jobs;
  for (int l = 0; l < kLayers; ++l)
    for (int mo = 0; mo < layerOut(l) / 16; ++mo)
      for (int mi = 0; mi < layerIn(l) / 16; ++mi) jobs.push_back(TileJob{l, mo, mi});
  nTileJobs_ = (uint32_t)jobs.size();
  hipOk(hipMalloc(&dTileJobs_, sizeof(TileJob) * jobs.size()), "hipMalloc");
  hipOk(hipMemcpy(dTileJobs_, jobs.data(), sizeof(TileJob) * jobs.size(), hipMemcpyHostToDevice), "copy");
  for (int c = 0; c < 3; ++c) boxLo_[c] = 0.0f, boxInv_[c] = 1.0f;
  std::vector<float> p(kParams, 0.0f);
  initParams(0, p.data());
  setParams(p.data());
}

Guide::~Guide() {
  for (void* q : {(void*)dParam_, (void*)dM_, (void*)dV_, (void*)dGrad_, (void*)dWT_, (void*)dW_, (void*)dLoss_,
                  (void*)dTileJobs_, (void*)dH_, (void*)dG_, (void*)dIn_, (void*)dRaw_, (void*)dShuffled_})
    if (q) (void)hipFree(q);
}

void Guide::initParams(uint32_t seed, float* p) {
  std::mt19937 rng(seed);
  for (int l = 0; l < kLayers; ++l) {
    int nIn = l == 0 ? 78 : kHidden, nOut = l == kLayers - 1 ? kOutUsed : kHidden;
    float lim = l < kLayers - 1 ? std::sqrt(6.0f / nIn) : 0.1f * std::sqrt(3.0f / nIn);
    std::uniform_real_distribution<float> U(-lim, lim);
    for (int o = 0; o < layerOut(l); ++o)
      for (int k = 0; k < layerIn(l); ++k) p[wOffset(l) + o * layerIn(l) + k] = (o < nOut && k < nIn) ? U(rng) : 0.0f;
    for (int o = 0; o < layerOut(l); ++o) p[bOffset(l) + o] = 0.0f;
  }
}

void Guide::setParams(const float* host) {
  hipOk(hipMemcpyAsync(dParam_, host, sizeof(float) * kParams, hipMemcpyHostToDevice, stream_), "copy");
  kZero<<<32, 256, 0, stream_>>>(dM_, kParams);
  kZero<<<32, 256, 0, stream_>>>(dV_, kParams);
  kZero<<<32, 256, 0, stream_>>>(dGrad_, kParams);
  kQuantize<<<64, 256, 0, stream_>>>(dParam_, dWT_, dW_);
  hipOk(hipGetLastError(), "guide init");
  hipOk(hipStreamSynchronize(stream_), "sync");
  step_ = 0;
}

void Guide::getParams(float* host) {
  hipOk(hipMemcpyAsync(host, dParam_, sizeof(float) * kParams, hipMemcpyDeviceToHost, stream_), "copy");
  hipOk(hipStreamSynchronize(stream_), "sync");
}

void Guide::setBox(const float lo[3], const float hi[3]) {
  for (int c = 0; c < 3; ++c) {
    float ext = hi[c] - lo[c];
    boxLo_[c] = lo[c];
    boxInv_[c] = ext > 0.0f ? 1.0f / ext : 1.0f;
  }
}

Net Guide::net() const {
  Net n;
  n.wT = dWT_;
  n.w = dW_;
  n.param = dParam_;
  for (int c = 0; c < 3; ++c) n.boxLo[c] = boxLo_[c], n.boxInv[c] = boxInv_[c];
  n.selLo = selLo_;
  n.selHi = selHi_;
  return n;
}

void Guide::setSelection(float lo, float hi) {
  selLo_ = lo;
  selHi_ = hi;
}

void Guide::inferDevice(const float4* dIn, uint32_t n, float4* dRaw) {
  if (n == 0) return;
  uint32_t tiles = (n + 15) / 16;
  uint32_t blocks = std::min<uint32_t>(768, (tiles + 3) / 4);  // about one resident round
  kInfer<<<blocks, 256, 0, stream_>>>(net(), dIn, n, dRaw);
  hipOk(hipGetLastError(), "kInfer");
}

void Guide::ensureTrain(uint32_t batch) {
  if (batch <= trainCap_) return;
  if (dH_) (void)hipFree(dH_);
  if (dG_) (void)hipFree(dG_);
  hipOk(hipMalloc(&dH_, sizeof(short) * (size_t)hOffset(kLayers) * batch), "hipMalloc");
  hipOk(hipMalloc(&dG_, sizeof(short) * (size_t)gOffset(kLayers) * batch), "hipMalloc");
  trainCap_ = batch;
}

void Guide::gradientAsync(const float4* dSamples, uint32_t n) {
  const uint32_t chunk = 512;
  uint32_t batch = ((n + chunk - 1) / chunk) * chunk;  // multiple of the GEMM chunk (and of 16)
  ensureTrain(batch);
  kZero<<<1, 64, 0, stream_>>>(dLoss_, 1);
  uint32_t tiles = batch / 16;
  uint32_t blocks = std::min<uint32_t>(2048, (tiles + 3) / 4);
  kTrainFwdBwd<<<blocks, 256, 0, stream_>>>(net(), dSamples, n, batch, 1.0f / (float)n, dH_, dG_, dLoss_);
  hipOk(hipGetLastError(), "kTrainFwdBwd");
  uint32_t waves = nTileJobs_ * (batch / chunk);
  kWeightGrad<<<(waves + 3) / 4, 256, 0, stream_>>>(dH_, dG_, batch, chunk, (const TileJob*)dTileJobs_, nTileJobs_,
                                                    dGrad_);
  hipOk(hipGetLastError(), "kWeightGrad");
}

float Guide::gradientDevice(const float4* dSamples, uint32_t n) {
  if (n == 0) return 0.0f;
  gradientAsync(dSamples, n);
  float loss = 0.0f;
  hipOk(hipMemcpyAsync(&loss, dLoss_, sizeof(float), hipMemcpyDeviceToHost, stream_), "copy");
  hipOk(hipStreamSynchronize(stream_), "sync");
  return loss / (float)n;
}

void Guide::adamStep(float lr) {
  step_ += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float c1 = 1.0f - (float)std::pow(0.9, step_), c2 = 1.0f - (float)std::pow(0.999, step_);
  kAdam<<<64, 256, 0, stream_>>>(dParam_, dM_, dV_, dGrad_, lr, b1, b2, eps, c1, c2, dWT_, dW_);
  hipOk(hipGetLastError(), "kAdam");
}

float Guide::weightSum(const float4* dSamples, uint32_t n) {
  if (n == 0) return 0.0f;
  kZero<<<1, 64, 0, stream_>>>(dLoss_, 1);
  kWeightSum<<<256, 256, 0, stream_>>>(dSamples, n, dLoss_);
  float sum = 0.0f;
  hipOk(hipMemcpyAsync(&sum, dLoss_, sizeof(float), hipMemcpyDeviceToHost, stream_), "copy");
  hipOk(hipStreamSynchronize(stream_), "sync");
  return sum;
}

void Guide::scaleWeights(float4* dSamples, uint32_t n, float factor) {
  if (n == 0) return;
  kWeightScale<<<256, 256, 0, stream_>>>(dSamples, n, factor, 3.0e38f);
  hipOk(hipGetLastError(), "kWeightScale");
}

// Scales the weights to mean 1 and caps them; false if they sum to zero.
float Guide::normaliseWeights(float4* dSamples, uint32_t n, float weightCap) {
  const float sum = weightSum(dSamples, n);
  if (!(sum > 0.0f)) return 0.0f;
  kWeightScale<<<256, 256, 0, stream_>>>(dSamples, n, (float)n / sum, weightCap);
  return sum;
}

float Guide::trainOnDevice(float4* dSamples, uint32_t n, float lr, float weightCap) {
  if (n == 0 || !(normaliseWeights(dSamples, n, weightCap) > 0.0f)) return 0.0f;
  zeroGrad();
  float loss = gradientDevice(dSamples, n);
  adamStep(lr);
  return loss;
}

float Guide::trainEpoch(float4* dSamples, uint32_t n, float lr, float weightCap, uint32_t minibatch,
                        uint32_t minSteps) {
  if (n == 0 || minibatch == 0 || !(normaliseWeights(dSamples, n, weightCap) > 0.0f)) return 0.0f;
  if (shuffledCap_ < n) {
    if (dShuffled_) (void)hipFree(dShuffled_);
    dShuffled_ = nullptr;
    hipOk(hipMalloc(&dShuffled_, sizeof(float4) * 4 * (size_t)n), "hipMalloc");
    shuffledCap_ = n;
  }
  auto gcd = [](uint32_t x, uint32_t y) {
    while (y) {
      uint32_t t = x % y;
      x = y;
      y = t;
    }
    return x;
  };
  const uint32_t perPass = (n + minibatch - 1) / minibatch;
  const uint32_t steps = std::max(perPass, minSteps);
  float loss = 0.0f;
  for (uint32_t st = 0; st < steps; ++st) {
    const uint32_t k = st % perPass;
    if (k == 0) {  // a new pass: permutation i -> (a i + b) mod n, a coprime to n
      const uint32_t pass = st / perPass;
      uint32_t a = (uint32_t)((double)n * (0.6180339887 + 0.0731 * (pass % 5))) | 1u;
      while (gcd(a, n) != 1u) a += 2u;
      const uint32_t b = (uint32_t)(((uint64_t)step_ * 7919u) % n);
      kPermute<<<512, 256, 0, stream_>>>(dSamples, dShuffled_, n, a % n == 0u ? 1u : a % n, b);
      hipOk(hipGetLastError(), "kPermute");
    }
    const uint32_t off = k * minibatch, m = std::min(minibatch, n - off);
    if (st == 0) zeroGrad();  // each Adam step leaves the gradient cleared
    if (st + 1 == steps) {  // the last step reports its loss
      loss = gradientDevice(dShuffled_ + (size_t)off * 4, m);
    } else {
      gradientAsync(dShuffled_ + (size_t)off * 4, m);
    }
    adamStep(lr);
  }
  return loss;
}

void Guide::getGrad(float* host) {
  hipOk(hipMemcpyAsync(host, dGrad_, sizeof(float) * kParams, hipMemcpyDeviceToHost, stream_), "copy");
  hipOk(hipStreamSynchronize(stream_), "sync");
}
void Guide::zeroGrad() { kZero<<<32, 256, 0, stream_>>>(dGrad_, kParams); }

}  // namespace guide

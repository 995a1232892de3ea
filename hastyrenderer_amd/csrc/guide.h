// Learned path guide: a small MLP (BF16 weights on MFMA, FP32 accumulation)
// maps (position, normal, outgoing direction) to a mixture of 8 von
// Mises-Fisher lobes; trained online on the tracer's own radiance samples.
// No counterpart in the reference renderer.
//
// Network: 80 inputs (12-frequency sin/cos encoding of the position
// normalised to the scene box = 72, normal 3, outgoing direction 3, 2 zero
// pads) -> 64 -> 64 -> 64 -> 64 (ReLU) -> 48 raw outputs (41 used: per lobe
// mean direction 3, log concentration 1, logit 1; then the logit z of the
// probability c = lo + (hi - lo) sigmoid(z) of drawing a vertex's direction
// from the guide rather than from its BSDF).  Weights are kept as FP32
// master copies, as BF16 W^T [out][in] for the forward pass and BF16 W
// [in][out] for the backward pass.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "pt_device.h"

namespace guide {

constexpr int kFreq = 12;
constexpr int kInPad = 80;
constexpr int kHidden = 64;
constexpr int kLobes = 8;
constexpr int kOut = 40;      // mixture outputs
constexpr int kSelOut = 40;   // raw output of the selection logit
constexpr int kOutUsed = 41;  // outputs with weights (the rest are padding)
constexpr int kOutPad = 48;
constexpr int kLayers = 5;
constexpr float kCLo = -4.0f, kCHi = 9.0f;  // log-concentration clamp

__host__ __device__ constexpr int layerIn(int l) { return l == 0 ? kInPad : kHidden; }
__host__ __device__ constexpr int layerOut(int l) { return l == kLayers - 1 ? kOutPad : kHidden; }
// Offsets are loops, not recursions: a recursive call with a run-time layer
// index stays a real call on the device and gives the kernel a dynamic stack
// (scratch memory for every wave).
// flat parameter layout: per layer W^T (out x in) then b (out)
__host__ __device__ constexpr int wOffset(int l) {
  int off = 0;
  for (int k = 0; k < l; ++k) off += layerIn(k) * layerOut(k) + layerOut(k);
  return off;
}
__host__ __device__ constexpr int bOffset(int l) { return wOffset(l) + layerIn(l) * layerOut(l); }
constexpr int kParams = bOffset(kLayers - 1) + layerOut(kLayers - 1);
// offsets of the BF16 weight matrices (biases excluded)
__host__ __device__ constexpr int wqOffset(int l) {
  int off = 0;
  for (int k = 0; k < l; ++k) off += layerIn(k) * layerOut(k);
  return off;
}
constexpr int kWeights = wqOffset(kLayers - 1) + layerIn(kLayers - 1) * layerOut(kLayers - 1);
// rows of the training activations: layer inputs (H) and output gradients (G)
__host__ __device__ constexpr int hOffset(int l) {
  int off = 0;
  for (int k = 0; k < l; ++k) off += layerIn(k);
  return off;
}
__host__ __device__ constexpr int gOffset(int l) {
  int off = 0;
  for (int k = 0; k < l; ++k) off += layerOut(k);
  return off;
}

// Device view of the network (read by inference and training kernels).
struct Net {
  const short* wT;     // BF16 W^T, kWeights
  const short* w;      // BF16 W (transposed copies), kWeights
  const float* param;  // FP32 master parameters (biases read from here)
  float boxLo[3], boxInv[3];  // position normalisation: (x - lo) * inv
  // learned selection probability c in [selLo, selHi]; selLo == selHi
  // (both 0 by default) leaves the selection logit untrained
  float selLo, selHi;
};

__device__ inline float selectionOf(const Net& n, float z) {
  return n.selLo + (n.selHi - n.selLo) / (1.0f + __expf(-z));
}

// ------------------------------------------------------------ BF16 ------
__host__ __device__ inline short toBf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return (short)((u >> 16) | 0x40u);  // NaN stays NaN
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__host__ __device__ inline float fromBf16(short h) {
  uint32_t u = ((uint32_t)(uint16_t)h) << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// ----------------------------------------------------- input features ----
// Feature k of a query (x normalised to [0,1]^3): 24a + 2f -> sin(2^f pi x_a),
// 24a + 2f + 1 -> cos(2^f pi x_a); 72..74 normal; 75..77 wo; 78, 79 zero.
__device__ inline float feature(int k, hdev::V3 x, hdev::V3 n, hdev::V3 wo) {
  if (k < 72) {
    int a = k / 24, r = k - 24 * a, f = r >> 1;
    float xa = a == 0 ? x.x : (a == 1 ? x.y : x.z);
    float t = (float)(1u << f) * xa;
    return (r & 1) ? cospif(t) : sinpif(t);
  }
  if (k < 75) return k == 72 ? n.x : (k == 73 ? n.y : n.z);
  if (k < 78) return k == 75 ? wo.x : (k == 76 ? wo.y : wo.z);
  return 0.0f;
}

// --------------------------------------------------------- vMF mixture ----
struct Mixture {
  hdev::V3 mu[kLobes];
  float kappa[kLobes];
  float w[kLobes];
};

__device__ inline Mixture mixtureFromRaw(const float* raw) {
  Mixture m;
  float amax = -3.0e38f;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) amax = fmaxf(amax, raw[5 * k + 4]);
  float sum = 0.0f;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) {
    hdev::V3 r = hdev::v3(raw[5 * k], raw[5 * k + 1], raw[5 * k + 2]);
    float nr = sqrtf(hdev::dot(r, r));
    m.mu[k] = nr > 1e-12f ? r / nr : hdev::v3(0.0f, 0.0f, 1.0f);
    m.kappa[k] = __expf(fminf(fmaxf(raw[5 * k + 3], kCLo), kCHi));
    m.w[k] = __expf(raw[5 * k + 4] - amax);
    sum += m.w[k];
  }
#pragma unroll
  for (int k = 0; k < kLobes; ++k) m.w[k] /= sum;
  return m;
}

__device__ inline float vmfPdf(hdev::V3 mu, float kappa, hdev::V3 d) {
  float norm = kappa / (2.0f * 3.14159265358979323846f * -expm1f(-2.0f * kappa));
  return norm * __expf(kappa * (hdev::dot(mu, d) - 1.0f));
}

__device__ inline float mixturePdf(const Mixture& m, hdev::V3 d) {
  float p = 0.0f;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) p += m.w[k] * vmfPdf(m.mu[k], m.kappa[k], d);
  return p;
}

// Lobe by its weight (xi0), then Jakob's stable inversion about mu.
// The first lobe whose cumulative weight exceeds xi0 (the last one if none
// does) is picked by selects with constant indices, so the mixture stays in
// registers.
__device__ inline hdev::V3 mixtureSample(const Mixture& m, float xi0, float xi1, float xi2) {
  float kappa = m.kappa[kLobes - 1];
  hdev::V3 mu = m.mu[kLobes - 1];
  float c = 0.0f;
  bool found = false;
#pragma unroll
  for (int j = 0; j < kLobes - 1; ++j) {
    c += m.w[j];
    const bool take = !found && xi0 < c;
    kappa = take ? m.kappa[j] : kappa;
    mu.x = take ? m.mu[j].x : mu.x;
    mu.y = take ? m.mu[j].y : mu.y;
    mu.z = take ? m.mu[j].z : mu.z;
    found = found || take;
  }
  float wz = 1.0f + __logf(xi1 + (1.0f - xi1) * __expf(-2.0f * kappa)) / kappa;
  wz = fminf(1.0f, fmaxf(-1.0f, wz));
  float s = sqrtf(fmaxf(0.0f, 1.0f - wz * wz));
  float phi = 2.0f * 3.14159265358979323846f * xi2;
  hdev::V3 t = fabsf(mu.x) > 0.9f ? hdev::v3(0.0f, 1.0f, 0.0f) : hdev::v3(1.0f, 0.0f, 0.0f);
  hdev::V3 b1 = hdev::normalize(hdev::cross(mu, t));
  hdev::V3 b2 = hdev::cross(mu, b1);
  return hdev::normalize(b1 * (s * __cosf(phi)) + b2 * (s * __sinf(phi)) + mu * wz);
}

// Mixture table of a shading point (10 float4): lobe k as (v_k, b_k) with
// v_k = kappa_k mu_k and b_k = log(w_k kappa_k / (2 pi (1 - e^{-2 kappa_k})))
// - kappa_k, so that lobe k's density term is exp(v_k . d + b_k); then the
// eight weights w_k.  Same mixture as mixtureFromRaw.
__device__ inline float4 lobeEntry(const float* r5, float wk) {
  hdev::V3 mu = hdev::v3(r5[0], r5[1], r5[2]);
  const float nr = sqrtf(hdev::dot(mu, mu));
  // reciprocals by v_rcp_f32 (1 ulp) rather than IEEE divisions: sampling
  // and density both read the table, so they stay consistent
  mu = nr > 1e-12f ? mu * __builtin_amdgcn_rcpf(nr) : hdev::v3(0.0f, 0.0f, 1.0f);
  const float kappa = __expf(fminf(fmaxf(r5[3], kCLo), kCHi));
  const float norm = kappa * __builtin_amdgcn_rcpf(2.0f * 3.14159265358979323846f * -expm1f(-2.0f * kappa));
  return make_float4(mu.x * kappa, mu.y * kappa, mu.z * kappa, __logf(wk * norm) - kappa);
}
__device__ inline float mixturePdfTable(const float4* g, hdev::V3 d) {
  float p = 0.0f;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) {
    const float4 a = hdev::ldg4(g + k);
    p += __expf(a.x * d.x + a.y * d.y + a.z * d.z + a.w);
  }
  return p;
}
__device__ inline hdev::V3 mixtureSampleTable(const float4* g, float xi0, float xi1, float xi2) {
  const float4 w0 = hdev::ldg4(g + kLobes), w1 = hdev::ldg4(g + kLobes + 1);
  const float w[kLobes] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
  int k = kLobes - 1;
  float c = 0.0f;
  bool found = false;
#pragma unroll
  for (int j = 0; j < kLobes - 1; ++j) {
    c += w[j];
    const bool take = !found && xi0 < c;
    k = take ? j : k;
    found = found || take;
  }
  const float4 a = hdev::ldg4(g + k);
  const float kappa = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
  const hdev::V3 mu = hdev::v3(a.x, a.y, a.z) / kappa;
  float wz = 1.0f + __logf(xi1 + (1.0f - xi1) * __expf(-2.0f * kappa)) / kappa;
  wz = fminf(1.0f, fmaxf(-1.0f, wz));
  const float s = sqrtf(fmaxf(0.0f, 1.0f - wz * wz));
  const float phi = 2.0f * 3.14159265358979323846f * xi2;
  const hdev::V3 t = fabsf(mu.x) > 0.9f ? hdev::v3(0.0f, 1.0f, 0.0f) : hdev::v3(1.0f, 0.0f, 0.0f);
  const hdev::V3 b1 = hdev::normalize(hdev::cross(mu, t));
  const hdev::V3 b2 = hdev::cross(mu, b1);
  return hdev::normalize(b1 * (s * __cosf(phi)) + b2 * (s * __sinf(phi)) + mu * wz);
}

// ------------------------------------------------------------- MFMA MLP ----
// Every layer is computed transposed, H_out^T (out x 16 queries) =
// W^T (out x in) . H_in^T (in x 16), by 16x16x16 tiles of
// v_mfma_f32_16x16x16_bf16.  Operand lanes:
//   A (16 x 16, row i, column k): lane l holds i = l % 16, k = 4 (l / 16) + v
//   B (16 x 16, row k, column j): lane l holds j = l % 16, k = 4 (l / 16) + v
//   D (16 x 16, row i, column j): lane l holds j = l % 16, i = 4 (l / 16) + v
// so an output tile in D layout is, after bias/ReLU and rounding, already the
// B operand of the next layer for the matching k-step.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// Round to nearest even, as toBf16 (gfx950 converts pairs in one instruction).
__device__ inline bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 r;
  r.x = __builtin_bit_cast(short, (__bf16)a);
  r.y = __builtin_bit_cast(short, (__bf16)b);
  r.z = __builtin_bit_cast(short, (__bf16)c);
  r.w = __builtin_bit_cast(short, (__bf16)d);
  return r;
}
__device__ inline bf16x4 load4(const short* p) { return *reinterpret_cast<const bf16x4*>(p); }

// One layer, transposed: z[mt] (D layout) = W^T tile rows . x[ks] + b.
template <int KS, int MT>
__device__ inline void layerFwd(const short* wT, const float* bias, int din, const bf16x4* x, f32x4* z) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const short* row = wT + (size_t)(16 * mt + r) * din + 4 * q;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) acc = mfma(load4(row + 16 * ks), x[ks], acc);
    const float* b = bias + 16 * mt + 4 * q;
    acc.x += b[0];
    acc.y += b[1];
    acc.z += b[2];
    acc.w += b[3];
    z[mt] = acc;
  }
}
__device__ inline bf16x4 reluPack(f32x4 z) {
  return pack4(fmaxf(z.x, 0.0f), fmaxf(z.y, 0.0f), fmaxf(z.z, 0.0f), fmaxf(z.w, 0.0f));
}

// Input features of query (x, n, wo) for the 5 k-steps of layer 0 (B layout).
// A lane's 4 consecutive features k..k+3 (k a multiple of 4): below 72 they
// are the sine/cosine pairs of octaves f and f+1 of one axis, computed as
// sinpi / cospi of 2^f x_a (the scaling is exact, so no angle is rounded
// before the reduction).  DOUBLE_ANGLE (the path tracer's inference) takes
// octave f+1 from octave f instead: sin 2a = 2 sin a cos a and
// cos 2a = (cos a - sin a)(cos a + sin a), a few FP32 ulp from the direct
// values (an absolute 1e-7 near zero), below the BF16 rounding of the
// operands; training and the library's inference keep the direct form.
template <bool DOUBLE_ANGLE = false>
__device__ inline void inputOperands(hdev::V3 x, hdev::V3 n, hdev::V3 wo, bf16x4* in) {
  const int q = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int ks = 0; ks < 5; ++ks) {
    const int k = 16 * ks + 4 * q;
    if (k < 72) {
      const int a = k / 24, f = (k - 24 * a) >> 1;
      const float xa = a == 0 ? x.x : (a == 1 ? x.y : x.z);
      const float t = xa * (float)(1u << f);
      float s0, c0, s1, c1;
      sincospif(t, &s0, &c0);
      if (DOUBLE_ANGLE) {
        s1 = 2.0f * s0 * c0;
        c1 = (c0 - s0) * (c0 + s0);
      } else {
        sincospif(t + t, &s1, &c1);
      }
      in[ks] = pack4(s0, c0, s1, c1);
    } else {
      in[ks] = pack4(feature(k, x, n, wo), feature(k + 1, x, n, wo), feature(k + 2, x, n, wo),
                     feature(k + 3, x, n, wo));
    }
  }
}

// Forward through all layers: raw[3] (D layout of the 48 outputs); hidden
// operands and pre-activations kept when the caller wants them.
struct Acts {
  bf16x4 h[kLayers][5];  // layer inputs (B layout); h[0] has 5 k-steps, others 4
  f32x4 z[kLayers - 1][4];
};
__device__ inline void forwardAll(const Net& net, Acts& a, f32x4* raw) {
  layerFwd<5, 4>(net.wT + wqOffset(0), net.param + bOffset(0), kInPad, a.h[0], a.z[0]);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) a.h[1][mt] = reluPack(a.z[0][mt]);
  layerFwd<4, 4>(net.wT + wqOffset(1), net.param + bOffset(1), kHidden, a.h[1], a.z[1]);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) a.h[2][mt] = reluPack(a.z[1][mt]);
  layerFwd<4, 4>(net.wT + wqOffset(2), net.param + bOffset(2), kHidden, a.h[2], a.z[2]);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) a.h[3][mt] = reluPack(a.z[2][mt]);
  layerFwd<4, 4>(net.wT + wqOffset(3), net.param + bOffset(3), kHidden, a.h[3], a.z[3]);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) a.h[4][mt] = reluPack(a.z[3][mt]);
  layerFwd<4, 3>(net.wT + wqOffset(4), net.param + bOffset(4), kHidden, a.h[4], raw);
}

// ---------------------------------------------- LDS-staged inference ----
// Inference kernels copy the BF16 W^T and the biases into LDS once per block
// (42,176 B) and loop over many 16-query tiles, so that each MFMA's weight
// operand is an LDS read instead of a global load per tile.  The copy is kept
// in operand order: the A operand of k-step ks of output tile mt of layer l
// is 64 lanes x 4 BF16 at wqOffset(l) + 4 (64 (KS mt + ks) + lane), so a
// wave's read is 512 contiguous bytes (rows of W^T 128 B apart would put the
// 16 lanes of a lane group in one bank).
constexpr int kBiasLds = 4 * kHidden + kOutPad;
__host__ __device__ constexpr int biasLds(int l) { return l * kHidden; }
struct LdsNet {
  short wT[kWeights];
  float bias[kBiasLds];
};
// Output row of the last layer held at LDS row p when the rows are grouped by
// lobe: lane group q of a 16-query tile holds rows 16 mt + 4 q + v (mt < 3,
// v < 4) of the output; with this order its 12 values are raw outputs
// 10 q .. 10 q + 9 (lobes 2q and 2q + 1) then outputs 40 + 2q and 41 + 2q
// (for q = 0 the selection logit, in the third value of tile 2).
__device__ inline int lobeMajorRow(int p) {
  const int mt = p >> 4, q = (p >> 2) & 3, v = p & 3, sl = 4 * mt + v;
  return sl < 10 ? 10 * q + sl : kOut + 2 * q + (sl - 10);
}
// lobeMajor: the last layer's rows in lobeMajorRow order (kGuideInfer);
// otherwise the network's own order.
__device__ inline void stageNet(const Net& net, LdsNet* L, bool lobeMajor = false) {
  const uint2* src = reinterpret_cast<const uint2*>(net.wT);
  uint2* dst = reinterpret_cast<uint2*>(L->wT);
  for (int i = threadIdx.x; i < kWeights / 4; i += blockDim.x) {
    int l = 0;
    while (l < kLayers - 1 && i >= wqOffset(l + 1) / 4) ++l;
    const int din = layerIn(l), KS = din / 16;
    const int local = i - wqOffset(l) / 4, lane = local & 63, t = local >> 6;
    const int mt = t / KS, ks = t - KS * mt;
    int row = 16 * mt + (lane & 15);
    if (lobeMajor && l == kLayers - 1) row = lobeMajorRow(row);
    dst[i] = src[(wqOffset(l) + row * din + 16 * ks + 4 * (lane >> 4)) / 4];
  }
  for (int i = threadIdx.x; i < kBiasLds; i += blockDim.x) {
    const int l = i < 4 * kHidden ? i / kHidden : 4;
    int o = i - biasLds(l);
    if (lobeMajor && l == kLayers - 1) o = lobeMajorRow(o);
    L->bias[i] = net.param[bOffset(l) + o];
  }
  __syncthreads();
}
template <int KS, int MT>
__device__ inline void layerFwdLds(const LdsNet* L, int l, const bf16x4* x, f32x4* z) {
  const int lane = threadIdx.x & 63, q = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const short* a = L->wT + wqOffset(l) + 4 * (64 * KS * mt + lane);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) acc = mfma(load4(a + 256 * ks), x[ks], acc);
    const float* b = L->bias + biasLds(l) + 16 * mt + 4 * q;
    acc.x += b[0];
    acc.y += b[1];
    acc.z += b[2];
    acc.w += b[3];
    z[mt] = acc;
  }
}
// Forward pass of one 16-query tile from the LDS copy (inputs in h0); only
// the raw outputs are kept.
__device__ inline void forwardLds(const LdsNet* L, const bf16x4* h0, f32x4* raw) {
  f32x4 z[4];
  bf16x4 h[4];
  layerFwdLds<5, 4>(L, 0, h0, z);
#pragma unroll
  for (int l = 1; l < kLayers - 1; ++l) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) h[mt] = reluPack(z[mt]);
    layerFwdLds<4, 4>(L, l, h, z);
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) h[mt] = reluPack(z[mt]);
  layerFwdLds<4, 3>(L, kLayers - 1, h, raw);
}

// Outputs of a 16-query tile for mixture tables (every lane of the wave
// calls this; lane l supplies query l % 16, position normalised; the LDS copy
// staged in lobeMajorRow order).  Lane group q = l / 16 receives its query's
// raw outputs 10 q .. 10 q + 9 (lobes 2q and 2q + 1) in r10, those lobes'
// mixture weights (the softmax over the 8 logits runs across the query's
// four lanes) in w0, w1, and in *logit raw output 40 (the selection logit)
// on lane group 0.  The inputs take the double-angle form of inputOperands.
__device__ inline void tileOutputs(const LdsNet* L, hdev::V3 x, hdev::V3 n, hdev::V3 wo, float* r10, float* w0,
                                   float* w1, float* logit) {
  bf16x4 h0[5];
  inputOperands<true>(x, n, wo, h0);
  f32x4 o[3];
  forwardLds(L, h0, o);
  r10[0] = o[0].x, r10[1] = o[0].y, r10[2] = o[0].z, r10[3] = o[0].w;
  r10[4] = o[1].x, r10[5] = o[1].y, r10[6] = o[1].z, r10[7] = o[1].w;
  r10[8] = o[2].x, r10[9] = o[2].y;
  float amax = fmaxf(r10[4], r10[9]);
  amax = fmaxf(amax, __shfl_xor(amax, 16));
  amax = fmaxf(amax, __shfl_xor(amax, 32));
  const float e0 = __expf(r10[4] - amax), e1 = __expf(r10[9] - amax);
  float sum = e0 + e1;
  sum += __shfl_xor(sum, 16);
  sum += __shfl_xor(sum, 32);
  // v_rcp_f32 (1 ulp) instead of two IEEE divisions: the weights then sum
  // to one within ~1e-7, and sampling and density read the same table
  const float inv = __builtin_amdgcn_rcpf(sum);
  *w0 = e0 * inv;
  *w1 = e1 * inv;
  *logit = o[2].z;
}

__device__ inline hdev::V3 normPos(const Net& net, hdev::V3 p) {
  return hdev::v3((p.x - net.boxLo[0]) * net.boxInv[0], (p.y - net.boxLo[1]) * net.boxInv[1],
            (p.z - net.boxLo[2]) * net.boxInv[2]);
}

// Raw outputs of one query per lane group: every lane of the wave calls this
// (MFMA is wave-wide); lane l supplies query l % 16 and receives outputs
// 16 mt + 4 (l / 16) + v (mt < 3, v < 4) of that query in raw[4 mt + v].
__device__ inline void inferOne(const Net& net, hdev::V3 x, hdev::V3 n, hdev::V3 wo, float* raw) {
  Acts a;
  inputOperands(normPos(net, x), n, wo, a.h[0]);
  f32x4 o[3];
  forwardAll(net, a, o);
#pragma unroll
  for (int mt = 0; mt < 3; ++mt) {
    raw[4 * mt] = o[mt].x;
    raw[4 * mt + 1] = o[mt].y;
    raw[4 * mt + 2] = o[mt].z;
    raw[4 * mt + 3] = o[mt].w;
  }
}

}  // namespace guide

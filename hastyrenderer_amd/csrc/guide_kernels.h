// Host side of the MLP path guide: parameters, Adam state, training buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "guide.h"

namespace guide {

class Guide {
 public:
  explicit Guide(hipStream_t stream);
  ~Guide();
  Guide(const Guide&) = delete;
  Guide& operator=(const Guide&) = delete;

  // He-uniform hidden layers, a small output layer, zero biases and padding.
  static void initParams(uint32_t seed, float* host);
  void setParams(const float* host);  // also resets Adam
  void getParams(float* host);
  void setBox(const float lo[3], const float hi[3]);
  // Range of the learned selection probability (lo < hi), or lo == hi for
  // an untrained selection output (the renderer's fixed fraction applies).
  void setSelection(float lo, float hi);
  bool learnedSelection() const { return selHi_ > selLo_; }
  Net net() const;

  // n queries (pos, normal, wo as 3 float4 each) -> 48 raw outputs each.
  void inferDevice(const float4* dIn, uint32_t n, float4* dRaw);
  // Adds d(mean loss)/d(params) of n samples (pos + F, normal, wo + the
  // density wi was drawn with or 0, wi + BSDF density of wi or 0) to the
  // gradient buffer; returns the mean loss.
  float gradientDevice(const float4* dSamples, uint32_t n);
  void adamStep(float lr);  // consumes and clears the gradient
  // One Adam step on n samples whose weights are first normalised to mean 1
  // and capped at weightCap (in place); returns the mean loss.
  float trainOnDevice(float4* dSamples, uint32_t n, float lr, float weightCap);
  // Adam steps of `minibatch` samples over n samples in pseudo-random order
  // (the samples arrive grouped by image region): one pass, or more passes
  // (each in another order) until minSteps steps are taken.  Weights are
  // normalised over all n first.  Returns the mean loss of the last step.
  float trainEpoch(float4* dSamples, uint32_t n, float lr, float weightCap, uint32_t minibatch,
                   uint32_t minSteps = 1);
  // Sum of n samples' weights; and F <- F * factor for n samples.
  float weightSum(const float4* dSamples, uint32_t n);
  void scaleWeights(float4* dSamples, uint32_t n, float factor);
  void getGrad(float* host);
  void zeroGrad();
  uint32_t steps() const { return step_; }
  hipStream_t stream() const { return stream_; }

 private:
  void ensureTrain(uint32_t batch);
  void gradientAsync(const float4* dSamples, uint32_t n);
  float normaliseWeights(float4* dSamples, uint32_t n, float weightCap);
  hipStream_t stream_;
  float* dParam_ = nullptr;
  float* dM_ = nullptr;
  float* dV_ = nullptr;
  float* dGrad_ = nullptr;
  short* dWT_ = nullptr;
  short* dW_ = nullptr;
  float* dLoss_ = nullptr;
  void* dTileJobs_ = nullptr;
  uint32_t nTileJobs_ = 0;
  short* dH_ = nullptr;
  short* dG_ = nullptr;
  uint32_t trainCap_ = 0;
  float4* dIn_ = nullptr;
  float4* dRaw_ = nullptr;
  float4* dShuffled_ = nullptr;
  uint32_t shuffledCap_ = 0;
  float boxLo_[3], boxInv_[3];
  float selLo_ = 0.0f, selHi_ = 0.0f;
  uint32_t step_ = 0;
};

}  // namespace guide

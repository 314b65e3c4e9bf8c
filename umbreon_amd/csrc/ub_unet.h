// ub_unet.h -- the learned denoiser: a U-Net over the GI irradiance field and
// its guides, with bf16 MFMA convolutions (fp32 accumulation) for inference
// and training on the device.
//
// Network (levels L, feature channels c[0..L-1], every count a multiple of
// 32): input 12 features (log(1 + E) rgb, albedo rgb, normal xyz, depth, and
// two noise descriptors: 1 / sqrt(samples) and the colour's standard error,
// zero where the producer has none) padded to 32 channels;
//   encoder level l: conv3x3(l == 0 ? input : avgpool2(e[l-1])) + ReLU,
//                    conv3x3 + ReLU -> e[l]
//   decoder level l = L-2 .. 0: conv3x3(concat(e[l], up2(d[l+1]))) + ReLU,
//                    conv3x3 + ReLU -> d[l]   (d[L-1] = e[L-1])
//   output: conv3x3(d[0]) -> 32 channels, of which the first 3 are added to
//   the input's log(1 + E): the predicted log(1 + E).
// Pooling, nearest upsampling and concatenation are folded into the
// convolutions' operand loads. Tensors are NHWC; activations bf16.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace ubg {

constexpr int kUnetMaxLevels = 6;
constexpr int kUnetFeatures = 12;  // input features per pixel
constexpr int kUnetInPad = 32;     // padded input channels
constexpr int kUnetOutPad = 32;    // padded output channels (3 used)

class Unet {
 public:
  ~Unet();
  // Param layout and device state; levels 2..6, channels multiples of 32 up
  // to 128. Returns 0 or sets err.
  int init(int levels, const int* channels, uint64_t seed, std::string& err);
  int64_t paramCount() const { return nParams_; }
  int levels() const { return L_; }
  // flat fp32 parameters: per convolution W[9][cout][cin] then b[cout]
  hipError_t getParams(float* host, hipStream_t st);
  hipError_t setParams(const float* host, hipStream_t st);
  // features: n x h x w x 10 fp32 (device); out: n x h x w x 3 fp32 log(1 + E)
  // (device). h and w must be multiples of 2^(levels - 1).
  hipError_t forward(int n, int h, int w, const float* feat, float* out, hipStream_t st);
  // Mean L1 (default) or mean squared loss of the prediction against target
  // (n x h x w x 3 log(1 + E), device) and the parameter gradients (device,
  // paramCount floats).
  hipError_t gradients(int n, int h, int w, const float* feat, const float* target, float* loss,
                       hipStream_t st);
  hipError_t copyGrads(float* host, hipStream_t st);
  // one Adam step (beta 0.9 / 0.999, eps 1e-8) with the last gradients
  hipError_t adam(float lr, hipStream_t st);
  // Convolution launch shape: a launch takes the widest pixel tiles per wave
  // (4 for 32 output channels, else 2) that still give it at least this many
  // workgroups (default 1,024); 0 = always the widest.
  void setMinWorkgroups(int64_t n) { minWgs_ = n; }
  // Training loss: false = mean L1, true = mean squared error.
  void setLossL2(bool l2) { lossL2_ = l2; }

 private:
  struct Src {
    int act = -1;   // activation index, -1 = none
    int mode = 0;   // 0 direct, 1 avgpool2 (source at twice the resolution), 2 nearest up2
  };
  struct Conv {
    int cin = 0, cout = 0, relu = 1, level = 0;
    Src s0, s1;
    int out = -1;          // activation index written (-1: the fp32 output buffer)
    int64_t wOff = 0, bOff = 0;
  };
  struct Act {
    int C = 0, level = 0;
  };
  hipError_t ensure(int n, int h, int w);
  void syncBf16(hipStream_t st);
  int L_ = 0;
  int ch_[kUnetMaxLevels] = {0};
  std::vector<Conv> convs_;
  std::vector<Act> acts_;
  int64_t nParams_ = 0;
  // device state
  float *params_ = nullptr, *grads_ = nullptr, *m_ = nullptr, *v_ = nullptr;
  uint16_t *wBf_ = nullptr, *wtBf_ = nullptr;  // bf16 weights and their tap-flipped transposes
  int64_t wCount_ = 0;                          // weights only (bf16 copies)
  std::vector<int64_t> wBfOff_;                 // per conv offset into wBf_ / wtBf_
  int step_ = 0;
  bool bf16Dirty_ = true;
  int64_t minWgs_ = 1024;
  bool lossL2_ = false;
  // activation scratch for the current shape
  int64_t capP0_ = 0;  // pixels (n h w) the activation and gradient buffers hold
  std::vector<uint16_t*> actBf_;   // bf16 activations
  std::vector<float*> actGrad_;    // fp32 gradients of the activations
  float* outF_ = nullptr;          // output conv, fp32, 32 channels
  float* dOut_ = nullptr;          // its gradient
  uint16_t* dz_ = nullptr;         // bf16 masked gradient of the conv being back-propagated
  float* dx_ = nullptr;            // fp32 gradient w.r.t. a conv's (transformed) input
  float* part_ = nullptr;          // split-K partials of the weight / bias gradients
  float* lossPart_ = nullptr;
  std::vector<void*> scratch_;
};

}  // namespace ubg

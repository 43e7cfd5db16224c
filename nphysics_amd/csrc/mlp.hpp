// nphysics_amd/csrc/mlp.hpp -- the warm-start network: 24 -> 128 -> 128 -> 3, GELU (tanh form),
// BF16 weights and activations on MFMA with FP32 accumulation.
//
// Parameter vector (FP32, the layout of nphys_mlp_get_params / nphys_warmstart_load):
//   W1[24][128] b1[128] W2[128][128] b2[128] W3[128][3] b3[3]      (20,099 trained parameters)
//   x_mean[24] x_scale[24] y_scale[3]                               (normalisation, from the data)
// A layer computes out[n] = sum_k in[k] W[k][n] + b[n]; inputs are (x - x_mean) / x_scale and the
// network's outputs times y_scale are the predicted impulses in the feature's impulse unit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nph {

constexpr int kFeat = 24;     // input features
constexpr int kHid = 128;     // hidden width
constexpr int kOut = 3;       // outputs
constexpr int kK1 = 32;       // layer-1 reduction padded to one MFMA k-block
constexpr int kN3 = 16;       // layer-3 outputs padded to one MFMA tile
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kFeat * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kOut;
constexpr int kNumTrained = kOffB3 + kOut;        // 20,099
constexpr int kOffXMean = kNumTrained;
constexpr int kOffXScale = kOffXMean + kFeat;
constexpr int kOffYScale = kOffXScale + kFeat;
constexpr int kNumParams = kOffYScale + kOut;     // 20,150

// BF16 operand images of the weights, as the MFMA kernels read them (row pitch padded by 8
// elements against LDS bank conflicts). "T" images are [out][in] (k contiguous) for the forward
// pass; plain images are [in][out] (out contiguous) for the backward pass.
constexpr int kP1 = kK1 + 8;   // W1T [128][40]
constexpr int kP2 = kHid + 8;  // W2T [128][136], W2 [128][136], W3T [16][136]
constexpr int kP3 = kK1 + 8;   // W3 [128][40] (k = outputs, padded to 32)
constexpr int kImgW1T = 0;
constexpr int kImgW2T = kImgW1T + kHid * kP1;
constexpr int kImgW3T = kImgW2T + kHid * kP2;
constexpr int kImgW2 = kImgW3T + kN3 * kP2;
constexpr int kImgW3 = kImgW2 + kHid * kP2;
constexpr int kImgElems = kImgW3 + kHid * kP3;  // bf16 elements
constexpr int kFwdImgElems = kImgW2;            // what the forward pass stages in LDS

// Device view of one network.
struct MlpDev {
  const uint16_t* img;   // [kImgElems] bf16 weight images
  const float* params;   // [kNumParams] FP32 (biases and normalisation are read from here)
};

// Training-time buffers of one batch of B rows (B a multiple of 64).
struct MlpBatch {
  int B;
  const float* X;        // [n][24] dataset features
  const float* T;        // [n][3] dataset targets (impulse units)
  const int* idx;        // [B] dataset rows of this batch
  uint16_t* A0;          // [B][32] normalised inputs, bf16
  float* Z1;             // [B][128] pre-activations
  float* Z2;
  uint16_t* H1;          // [B][128] activations, bf16
  uint16_t* H2;
  float* Y;              // [B][3] outputs (normalised units)
  uint16_t* D3;          // [B][32] dL/dY, bf16 (k padded)
  uint16_t* D2;          // [B][128] dL/dZ2, bf16
  uint16_t* D1;          // [B][128] dL/dZ1, bf16
  double* loss_part;     // [blocks] partial losses
  float* gpart;          // [blocks][kNumTrained] per-block weight-gradient partials
};

// Launchers (k_mlp.hip).
void mlp_pack_images(const float* params, uint16_t* img, hipStream_t s);
// Inference: out[r][3] = network(features[r]) in impulse units.
void mlp_infer(const MlpDev& m, const float* X, int n, float* out, hipStream_t s);
// One training pass over a batch: forward with stores, loss and backward, weight gradients
// reduced into grad[kNumTrained] (mean over the batch), loss (mean) into *loss.
int mlp_grad_blocks(int B);
void mlp_batch_gradients(const MlpDev& m, MlpBatch& b, float* grad, double* loss, hipStream_t s);
void mlp_adam(float* params, float* mom1, float* mom2, const float* grad, float lr, float beta1, float beta2,
              float eps, int t, uint16_t* img, hipStream_t s);

}  // namespace nph

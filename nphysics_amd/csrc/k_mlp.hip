// nphysics_amd/csrc/k_mlp.hip -- BF16 MFMA kernels of the warm-start network (mlp.hpp):
// inference, the training forward/backward passes, weight gradients and Adam.
//
// All matrix products use v_mfma_f32_16x16x32_bf16 (gfx950): a wave multiplies a 16x32 A block
// by a 32x16 B block into a 16x16 FP32 accumulator. Operand fragments: lane l holds 8 bf16 of
// row (l & 15) of A, and of column (l & 15) of B, at reduction indices 8 (l >> 4) .. +7 of the
// 32-block (the same mapping for A and B, so any order the hardware sums in is consistent); the
// accumulator gives lane l column (l & 15), rows 4 (l >> 4) .. +3.
#include <hip/hip_runtime.h>

#include "mlp.hpp"

namespace nph {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(uint4 a, uint4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                  0, 0);
}

// FP32 -> BF16, round to nearest even (inputs are finite).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ __forceinline__ uint4 pack8(const float* v) {
  uint4 r;
  r.x = uint32_t(f2bf(v[0])) | (uint32_t(f2bf(v[1])) << 16);
  r.y = uint32_t(f2bf(v[2])) | (uint32_t(f2bf(v[3])) << 16);
  r.z = uint32_t(f2bf(v[4])) | (uint32_t(f2bf(v[5])) << 16);
  r.w = uint32_t(f2bf(v[6])) | (uint32_t(f2bf(v[7])) << 16);
  return r;
}

// GELU, tanh form: 0.5 x (1 + tanh(c (x + a x^3))), c = sqrt(2/pi), a = 0.044715.
constexpr float kGc = 0.7978845608028654f;
constexpr float kGa = 0.044715f;
__device__ __forceinline__ float tanh_(float u) { return 1.0f - 2.0f / (__expf(2.0f * u) + 1.0f); }
__device__ __forceinline__ float gelu(float x) {
  const float t = tanh_(kGc * (x + kGa * x * x * x));
  return 0.5f * x * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float t = tanh_(kGc * (x + kGa * x * x * x));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * kGc * (1.0f + 3.0f * kGa * x * x);
}

__device__ __forceinline__ uint4 lds16(const uint16_t* p) { return *reinterpret_cast<const uint4*>(p); }

constexpr int kWaves = 4;
constexpr int kRowsPerBlock = 16 * kWaves;

// ------------------------------------------------------------------------------ forward pass
// One wave: 16 rows through the three layers. Weights (transposed images) and biases in LDS;
// the layer activations pass between layers through the wave's 16 x kP2 LDS tile.
template <bool kTrain>
__global__ void __launch_bounds__(256) k_mlp_fwd(MlpDev m, const float* __restrict__ X, const int* __restrict__ idx,
                                                 int n, float* __restrict__ out, MlpBatch b) {
  __shared__ __align__(16) uint16_t s_w[kFwdImgElems];
  __shared__ __align__(16) uint16_t s_h[kWaves][16 * kP2];
  __shared__ float s_b1[kHid], s_b2[kHid], s_b3[kN3];
  __shared__ float s_mean[kFeat], s_scale[kFeat], s_ys[kOut];
  const int tid = threadIdx.x;
  for (int i = tid; i < kFwdImgElems / 8; i += blockDim.x)
    reinterpret_cast<uint4*>(s_w)[i] = reinterpret_cast<const uint4*>(m.img)[i];
  for (int i = tid; i < kHid; i += blockDim.x) {
    s_b1[i] = m.params[kOffB1 + i];
    s_b2[i] = m.params[kOffB2 + i];
  }
  if (tid < kN3) s_b3[tid] = tid < kOut ? m.params[kOffB3 + tid] : 0.0f;
  if (tid < kFeat) {
    s_mean[tid] = m.params[kOffXMean + tid];
    s_scale[tid] = m.params[kOffXScale + tid];
  }
  if (tid < kOut) s_ys[tid] = m.params[kOffYScale + tid];
  __syncthreads();

  const int wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, kq = lane >> 4;
  uint16_t* H = s_h[wave];
  for (int base = blockIdx.x * kRowsPerBlock; base < n; base += gridDim.x * kRowsPerBlock) {
    const int r0 = base + 16 * wave;
    // layer 1: A = normalised features of row r0 + li, k = 8 kq .. + 7 (zero beyond 24)
    uint4 a1;
    {
      const int r = r0 + li;
      float v[8];
      const int row = r < n ? (kTrain ? idx[r] : r) : -1;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int k = 8 * kq + t;
        v[t] = (row >= 0 && k < kFeat) ? (X[long(row) * kFeat + k] - s_mean[k]) / s_scale[k] : 0.0f;
      }
      a1 = pack8(v);
      if (kTrain && r < n) *reinterpret_cast<uint4*>(b.A0 + long(r) * kK1 + 8 * kq) = a1;
    }
#pragma unroll
    for (int nt = 0; nt < kHid / 16; ++nt) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      acc = mfma(a1, lds16(s_w + kImgW1T + (16 * nt + li) * kP1 + 8 * kq), acc);
      const int col = 16 * nt + li;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = 4 * kq + q;
        const float z = acc[q] + s_b1[col];
        const uint16_t h = f2bf(gelu(z));
        H[row * kP2 + col] = h;
        if (kTrain && r0 + row < n) {
          b.Z1[long(r0 + row) * kHid + col] = z;
          b.H1[long(r0 + row) * kHid + col] = h;
        }
      }
    }
    __syncthreads();
    // layer 2
    uint4 a2[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) a2[kb] = lds16(H + li * kP2 + 32 * kb + 8 * kq);
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < kHid / 16; ++nt) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        acc = mfma(a2[kb], lds16(s_w + kImgW2T + (16 * nt + li) * kP2 + 32 * kb + 8 * kq), acc);
      const int col = 16 * nt + li;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = 4 * kq + q;
        const float z = acc[q] + s_b2[col];
        const uint16_t h = f2bf(gelu(z));
        H[row * kP2 + col] = h;
        if (kTrain && r0 + row < n) {
          b.Z2[long(r0 + row) * kHid + col] = z;
          b.H2[long(r0 + row) * kHid + col] = h;
        }
      }
    }
    __syncthreads();
    // layer 3 (outputs padded to 16)
    {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        acc = mfma(lds16(H + li * kP2 + 32 * kb + 8 * kq), lds16(s_w + kImgW3T + li * kP2 + 32 * kb + 8 * kq), acc);
      if (li < kOut) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = r0 + 4 * kq + q;
          if (r >= n) continue;
          const float y = acc[q] + s_b3[li];
          if (kTrain) b.Y[long(r) * kOut + li] = y;
          else out[long(r) * kOut + li] = y * s_ys[li];
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------ backward pass
// One wave: 16 rows. dL/dY of the mean squared error in normalised units, then
// dZ2 = (dY W3^T) * gelu'(Z2) and dZ1 = (dZ2 W2^T) * gelu'(Z1); the bf16 results are stored for
// the weight-gradient reduction. Per-block loss partials in double.
__global__ void __launch_bounds__(256) k_mlp_bwd(MlpDev m, MlpBatch b) {
  __shared__ __align__(16) uint16_t s_w2[kHid * kP2];
  __shared__ __align__(16) uint16_t s_w3[kHid * kP3];
  __shared__ __align__(16) uint16_t s_t[kWaves][16 * kP2];
  __shared__ float s_ys[kOut];
  __shared__ double s_loss[256];
  const int tid = threadIdx.x;
  for (int i = tid; i < kHid * kP2 / 8; i += blockDim.x)
    reinterpret_cast<uint4*>(s_w2)[i] = reinterpret_cast<const uint4*>(m.img + kImgW2)[i];
  for (int i = tid; i < kHid * kP3 / 8; i += blockDim.x)
    reinterpret_cast<uint4*>(s_w3)[i] = reinterpret_cast<const uint4*>(m.img + kImgW3)[i];
  if (tid < kOut) s_ys[tid] = m.params[kOffYScale + tid];
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, kq = lane >> 4;
  uint16_t* Tt = s_t[wave];
  const int n = b.B;
  const float inv_b = 1.0f / float(n);
  double loss = 0.0;
  for (int base = blockIdx.x * kRowsPerBlock; base < n; base += gridDim.x * kRowsPerBlock) {
    const int r0 = base + 16 * wave;
    // dY tile [16][32] in the wave's LDS tile (columns >= 3 zero)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = 4 * kq + q, r = r0 + row;
      float d = 0.0f;
      if (li < kOut && r < n) {
        const float e = b.Y[long(r) * kOut + li] - b.T[long(b.idx[r]) * kOut + li] / s_ys[li];
        loss += double(e) * double(e);
        d = 2.0f * e * inv_b;
      }
      const uint16_t h = f2bf(d);
      Tt[row * kP2 + li] = h;
      Tt[row * kP2 + 16 + li] = 0;
      if (r < n) {
        b.D3[long(r) * kK1 + li] = h;
        b.D3[long(r) * kK1 + 16 + li] = 0;
      }
    }
    __syncthreads();
    const uint4 ay = lds16(Tt + li * kP2 + 8 * kq);
    __syncthreads();
    // dZ2 = (dY W3^T) * gelu'(Z2): B operand W3 image [hidden][k = output]
#pragma unroll
    for (int nt = 0; nt < kHid / 16; ++nt) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      acc = mfma(ay, lds16(s_w3 + (16 * nt + li) * kP3 + 8 * kq), acc);
      const int col = 16 * nt + li;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = 4 * kq + q, r = r0 + row;
        const float z = r < n ? b.Z2[long(r) * kHid + col] : 0.0f;
        const uint16_t h = f2bf(acc[q] * gelu_grad(z));
        Tt[row * kP2 + col] = h;
        if (r < n) b.D2[long(r) * kHid + col] = h;
      }
    }
    __syncthreads();
    uint4 a2[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) a2[kb] = lds16(Tt + li * kP2 + 32 * kb + 8 * kq);
    // dZ1 = (dZ2 W2^T) * gelu'(Z1): B operand W2 image [layer-1 unit][k = layer-2 unit]
#pragma unroll
    for (int nt = 0; nt < kHid / 16; ++nt) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        acc = mfma(a2[kb], lds16(s_w2 + (16 * nt + li) * kP2 + 32 * kb + 8 * kq), acc);
      const int col = 16 * nt + li;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = r0 + 4 * kq + q;
        if (r < n) b.D1[long(r) * kHid + col] = f2bf(acc[q] * gelu_grad(b.Z1[long(r) * kHid + col]));
      }
    }
    __syncthreads();
  }
  s_loss[tid] = loss;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off) s_loss[tid] += s_loss[tid + off];
    __syncthreads();
  }
  if (tid == 0) b.loss_part[blockIdx.x] = s_loss[0] * double(inv_b);
}

// ------------------------------------------------------------------------------ weight gradients
// dW = A^T D over the batch rows of this block, for (A, D) = (A0, D1) [32 x 128], (H1, D2)
// [128 x 128], (H2, D3) [128 x 16]; bias gradients as ones^T D. Rows are staged 32 at a time,
// transposed into LDS (batch index contiguous) so both operands are 16-byte fragment loads.
constexpr int kStage = 32;
constexpr int kTP = kStage + 8;  // pitch of a transposed column
// LDS columns: A0 [0,32) H1 [32,160) H2 [160,288) D1 [288,416) D2 [416,544) D3 [544,560)
constexpr int kColA0 = 0, kColH1 = 32, kColH2 = 160, kColD1 = 288, kColD2 = 416, kColD3 = 544, kCols = 560;
// Output tiles: dW1 2x8, dW2 8x8, dW3 8x1, db1 8, db2 8, db3 1.
constexpr int kTiles = 16 + 64 + 8 + 17;

__device__ __forceinline__ void tile_of(int t, int& acol, int& dcol, int& kind, int& out) {
  // kind 0: weight tile (out = parameter offset of element (0, 0) and row stride in out>>20)
  if (t < 16) {
    acol = kColA0 + 16 * (t / 8);
    dcol = kColD1 + 16 * (t % 8);
    kind = 1;
  } else if (t < 80) {
    t -= 16;
    acol = kColH1 + 16 * (t / 8);
    dcol = kColD2 + 16 * (t % 8);
    kind = 2;
  } else if (t < 88) {
    t -= 80;
    acol = kColH2 + 16 * t;
    dcol = kColD3;
    kind = 3;
  } else if (t < 96) {
    acol = -1;
    dcol = kColD1 + 16 * (t - 88);
    kind = 4;
  } else if (t < 104) {
    acol = -1;
    dcol = kColD2 + 16 * (t - 96);
    kind = 5;
  } else {
    acol = -1;
    dcol = kColD3;
    kind = 6;
  }
  out = t;
}

__global__ void __launch_bounds__(256) k_mlp_wgrad(MlpBatch b, int rows_per_block) {
  __shared__ __align__(16) uint16_t s_c[kCols * kTP];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, kq = lane >> 4;
  constexpr int kPerWave = (kTiles + kWaves - 1) / kWaves;
  f32x4 acc[kPerWave];
#pragma unroll
  for (int i = 0; i < kPerWave; ++i) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const uint4 ones = make_uint4(0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u);
  const int r_begin = blockIdx.x * rows_per_block;
  const int r_end = min(b.B, r_begin + rows_per_block);
  for (int s0 = r_begin; s0 < r_end; s0 += kStage) {
    // stage 32 rows: each thread moves 8 consecutive columns of one row (16 bytes)
    for (int w = tid; w < kStage * (kCols / 8); w += blockDim.x) {
      const int row = w / (kCols / 8), c8 = 8 * (w % (kCols / 8));
      const int r = s0 + row;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (r < r_end) {
        const uint16_t* src;
        if (c8 < kColH1) src = b.A0 + long(r) * kK1 + (c8 - kColA0);
        else if (c8 < kColH2) src = b.H1 + long(r) * kHid + (c8 - kColH1);
        else if (c8 < kColD1) src = b.H2 + long(r) * kHid + (c8 - kColH2);
        else if (c8 < kColD2) src = b.D1 + long(r) * kHid + (c8 - kColD1);
        else if (c8 < kColD3) src = b.D2 + long(r) * kHid + (c8 - kColD2);
        else src = b.D3 + long(r) * kK1 + (c8 - kColD3);
        v = *reinterpret_cast<const uint4*>(src);
      }
      const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        s_c[(c8 + 2 * t) * kTP + row] = uint16_t(u[t] & 0xffffu);
        s_c[(c8 + 2 * t + 1) * kTP + row] = uint16_t(u[t] >> 16);
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPerWave; ++i) {
      const int t = wave * kPerWave + i;
      if (t >= kTiles) break;
      int acol, dcol, kind, o;
      tile_of(t, acol, dcol, kind, o);
      const uint4 a = acol >= 0 ? lds16(s_c + (acol + li) * kTP + 8 * kq) : ones;
      const uint4 d = lds16(s_c + (dcol + li) * kTP + 8 * kq);
      acc[i] = mfma(a, d, acc[i]);
    }
    __syncthreads();
  }
  // write this block's partial gradients: tile rows 4 kq .. +3, column li
  float* g = b.gpart + long(blockIdx.x) * kNumTrained;
#pragma unroll
  for (int i = 0; i < kPerWave; ++i) {
    const int t = wave * kPerWave + i;
    if (t >= kTiles) break;
    int acol, dcol, kind, o;
    tile_of(t, acol, dcol, kind, o);
    const int ncol = dcol + li;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int arow = acol + 4 * kq + q;
      const float v = acc[i][q];
      if (kind == 1) {
        const int k = arow - kColA0, nn = ncol - kColD1;
        if (k < kFeat) g[kOffW1 + k * kHid + nn] = v;
      } else if (kind == 2) {
        g[kOffW2 + (arow - kColH1) * kHid + (ncol - kColD2)] = v;
      } else if (kind == 3) {
        const int nn = ncol - kColD3;
        if (nn < kOut) g[kOffW3 + (arow - kColH2) * kOut + nn] = v;
      } else if (4 * kq + q == 0) {
        if (kind == 4) g[kOffB1 + ncol - kColD1] = v;
        else if (kind == 5) g[kOffB2 + ncol - kColD2] = v;
        else if (ncol - kColD3 < kOut) g[kOffB3 + ncol - kColD3] = v;
      }
    }
  }
}

// Sum of the block partials in block order (deterministic), and of the loss partials.
__global__ void k_mlp_reduce(const float* __restrict__ gpart, int nblocks, float* __restrict__ grad,
                             const double* __restrict__ loss_part, int nloss, double* loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kNumTrained) {
    float s = 0.0f;
    for (int k = 0; k < nblocks; ++k) s += gpart[long(k) * kNumTrained + i];
    grad[i] = s;
  }
  if (i == 0) {
    double l = 0.0;
    for (int k = 0; k < nloss; ++k) l += loss_part[k];
    *loss = l;
  }
}

__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ m1, float* __restrict__ m2,
                           const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNumTrained) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  p[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// BF16 images of the weights (layouts in mlp.hpp).
__global__ void k_mlp_pack(const float* __restrict__ p, uint16_t* __restrict__ img) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kImgElems) return;
  float v = 0.0f;
  if (e < kImgW2T) {
    const int nn = (e - kImgW1T) / kP1, k = (e - kImgW1T) % kP1;
    if (k < kFeat) v = p[kOffW1 + k * kHid + nn];
  } else if (e < kImgW3T) {
    const int nn = (e - kImgW2T) / kP2, k = (e - kImgW2T) % kP2;
    if (k < kHid) v = p[kOffW2 + k * kHid + nn];
  } else if (e < kImgW2) {
    const int nn = (e - kImgW3T) / kP2, k = (e - kImgW3T) % kP2;
    if (nn < kOut && k < kHid) v = p[kOffW3 + k * kOut + nn];
  } else if (e < kImgW3) {
    const int nn = (e - kImgW2) / kP2, k = (e - kImgW2) % kP2;
    if (k < kHid) v = p[kOffW2 + nn * kHid + k];
  } else {
    const int nn = (e - kImgW3) / kP3, k = (e - kImgW3) % kP3;
    if (k < kOut) v = p[kOffW3 + nn * kOut + k];
  }
  img[e] = f2bf(v);
}

inline int fwd_blocks(int n) {
  const int need = (n + kRowsPerBlock - 1) / kRowsPerBlock;
  return need < 1 ? 1 : (need > 1024 ? 1024 : need);
}

}  // namespace

void mlp_pack_images(const float* params, uint16_t* img, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((kImgElems + 255) / 256), dim3(256), 0, s, params, img);
}

void mlp_infer(const MlpDev& m, const float* X, int n, float* out, hipStream_t s) {
  if (n <= 0) return;
  MlpBatch none{};
  hipLaunchKernelGGL(k_mlp_fwd<false>, dim3(fwd_blocks(n)), dim3(256), 0, s, m, X, nullptr, n, out, none);
}

int mlp_grad_blocks(int B) { return (B + 255) / 256; }

void mlp_batch_gradients(const MlpDev& m, MlpBatch& b, float* grad, double* loss, hipStream_t s) {
  const int nfb = fwd_blocks(b.B);
  hipLaunchKernelGGL(k_mlp_fwd<true>, dim3(nfb), dim3(256), 0, s, m, b.X, b.idx, b.B, nullptr, b);
  hipLaunchKernelGGL(k_mlp_bwd, dim3(nfb), dim3(256), 0, s, m, b);
  const int gb = mlp_grad_blocks(b.B);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(gb), dim3(256), 0, s, b, 256);
  // loss partials of the backward kernel: nfb blocks
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNumTrained + 255) / 256), dim3(256), 0, s, b.gpart, gb, grad,
                     b.loss_part, nfb, loss);
}

void mlp_adam(float* params, float* mom1, float* mom2, const float* grad, float lr, float beta1, float beta2,
              float eps, int t, uint16_t* img, hipStream_t s) {
  const float c1 = 1.0f - powf(beta1, float(t));
  const float c2 = 1.0f - powf(beta2, float(t));
  hipLaunchKernelGGL(k_mlp_adam, dim3((kNumTrained + 255) / 256), dim3(256), 0, s, params, mom1, mom2, grad, lr,
                     beta1, beta2, eps, c1, c2);
  mlp_pack_images(params, img, s);
}

}  // namespace nph

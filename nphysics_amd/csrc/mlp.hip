// nphysics_amd/csrc/mlp.hip -- trainer of the warm-start network and its C ABI (nphys_mlp_*).
//
// The network and its training have no reference counterpart: the reference warm-starts only
// from the impulse cache (signorini_coulomb_pyramid_model.rs:104-108). Training runs on the GPU:
// the dataset lives in HBM, each mini-batch is one forward pass with stores, one backward pass,
// a weight-gradient reduction and an Adam update, all on the trainer's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/nphys.h"
#include "common.hpp"
#include "mlp.hpp"

using namespace nph;

struct nphys_mlp {
  hipStream_t stream = nullptr;
  DBuf params, mom1, mom2, grad, img;
  DBuf X, T, perm, losses;
  DBuf A0, Z1, Z2, H1, H2, Y, D3, D2, D1, loss_part, gpart, fx, fy;
  long n = 0;          // dataset rows
  int bcap = 0;        // batch capacity of the buffers
  long adam_t = 0;     // Adam steps taken
  uint64_t seed = 1;

  ~nphys_mlp() {
    if (stream) (void)hipStreamDestroy(stream);
  }

  MlpDev dev() const { return MlpDev{img.as<uint16_t>(), params.as<float>()}; }

  hipError_t ensure_batch(int B) {
    if (B <= bcap) return hipSuccess;
    const size_t b = size_t(B);
    hipError_t e;
    if ((e = A0.ensure(b * kK1 * 2)) || (e = Z1.ensure(b * kHid * 4)) || (e = Z2.ensure(b * kHid * 4)) ||
        (e = H1.ensure(b * kHid * 2)) || (e = H2.ensure(b * kHid * 2)) || (e = Y.ensure(b * kOut * 4)) ||
        (e = D3.ensure(b * kK1 * 2)) || (e = D2.ensure(b * kHid * 2)) || (e = D1.ensure(b * kHid * 2)) ||
        (e = loss_part.ensure(size_t(1024) * 8)) ||
        (e = gpart.ensure(size_t(mlp_grad_blocks(B)) * kNumTrained * 4)))
      return e;
    bcap = B;
    return hipSuccess;
  }

  MlpBatch batch(int B, const float* Xd, const float* Td, const int* idx) {
    MlpBatch b;
    b.B = B;
    b.X = Xd;
    b.T = Td;
    b.idx = idx;
    b.A0 = A0.as<uint16_t>();
    b.Z1 = Z1.as<float>();
    b.Z2 = Z2.as<float>();
    b.H1 = H1.as<uint16_t>();
    b.H2 = H2.as<uint16_t>();
    b.Y = Y.as<float>();
    b.D3 = D3.as<uint16_t>();
    b.D2 = D2.as<uint16_t>();
    b.D1 = D1.as<uint16_t>();
    b.loss_part = loss_part.as<double>();
    b.gpart = gpart.as<float>();
    return b;
  }
};

namespace {

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double uniform01(uint64_t& s) { return double(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

// He-normal weights (Box-Muller over SplitMix64), zero biases, identity normalisation.
void init_params(uint64_t seed, std::vector<float>& p) {
  p.assign(kNumParams, 0.0f);
  uint64_t s = seed;
  auto normal = [&]() {
    const double u1 = std::max(uniform01(s), 1e-300), u2 = uniform01(s);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  for (int i = 0; i < kFeat * kHid; ++i) p[kOffW1 + i] = float(normal() * std::sqrt(2.0 / kFeat));
  for (int i = 0; i < kHid * kHid; ++i) p[kOffW2 + i] = float(normal() * std::sqrt(2.0 / kHid));
  for (int i = 0; i < kHid * kOut; ++i) p[kOffW3 + i] = float(normal() * std::sqrt(1.0 / kHid));
  for (int i = 0; i < kFeat; ++i) p[kOffXScale + i] = 1.0f;
  for (int i = 0; i < kOut; ++i) p[kOffYScale + i] = 1.0f;
}

// Scratch of nphys_mlp_gradients_noalloc: the training-pass buffers of one batch of at most
// kNoAllocMaxRows rows (what ensure_batch allocates for nphys_mlp_gradients) and the identity row
// index, carved out of one file-scope device array at 256-byte aligned offsets (about 9.9 MB; every
// device the library runs on holds its copy until the process ends). Every byte a call reads is
// written earlier in the same call, so nothing needs clearing between calls.
constexpr int64_t kNoAllocMaxRows = 4096;
constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t kR = size_t(kNoAllocMaxRows);
constexpr size_t kScrIdx = 0;
constexpr size_t kScrA0 = align256(kScrIdx + kR * 4);
constexpr size_t kScrZ1 = align256(kScrA0 + kR * kK1 * 2);
constexpr size_t kScrZ2 = align256(kScrZ1 + kR * kHid * 4);
constexpr size_t kScrH1 = align256(kScrZ2 + kR * kHid * 4);
constexpr size_t kScrH2 = align256(kScrH1 + kR * kHid * 2);
constexpr size_t kScrY = align256(kScrH2 + kR * kHid * 2);
constexpr size_t kScrD3 = align256(kScrY + kR * kOut * 4);
constexpr size_t kScrD2 = align256(kScrD3 + kR * kK1 * 2);
constexpr size_t kScrD1 = align256(kScrD2 + kR * kHid * 2);
constexpr size_t kScrLoss = align256(kScrD1 + kR * kHid * 2);
constexpr size_t kScrGpart = align256(kScrLoss + size_t(1024) * 8);
constexpr size_t kScratchBytes = align256(kScrGpart + ((kR + 255) / 256) * size_t(kNumTrained) * 4);

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Device address of `scratch` on the current device, looked up once per device (with that device
// current) and kept: host code uses this address, never the array's name.
int scratch_on_current_device(unsigned char** out) {
  int dev = 0;
  HIPCK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices) return set_err(NPHYS_ERR_CAPACITY, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIPCK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return set_err(NPHYS_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[dev] = ptr;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return NPHYS_OK;
}

__global__ void k_mlp_iota(int* __restrict__ idx, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) idx[i] = i;
}

}  // namespace

extern "C" {

int nphys_mlp_create(uint64_t seed, nphys_mlp** out) {
  if (!out) return set_err(NPHYS_ERR_INVALID, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(NPHYS_ERR_NO_DEVICE, "no HIP device");
  nphys_mlp* m = new nphys_mlp();
  m->seed = seed;
  std::vector<float> p;
  init_params(seed, p);
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) ||
      (e = m->params.ensure(kNumParams * 4)) || (e = m->mom1.ensure(kNumTrained * 4)) ||
      (e = m->mom2.ensure(kNumTrained * 4)) || (e = m->grad.ensure(kNumTrained * 4)) ||
      (e = m->img.ensure(size_t(kImgElems) * 2)) || (e = m->fy.ensure(4)) ||
      (e = hipMemcpy(m->params.p, p.data(), kNumParams * 4, hipMemcpyHostToDevice)) ||
      (e = hipMemset(m->mom1.p, 0, kNumTrained * 4)) || (e = hipMemset(m->mom2.p, 0, kNumTrained * 4))) {
    delete m;
    return hip_err(e);
  }
  mlp_pack_images(m->params.as<float>(), m->img.as<uint16_t>(), m->stream);
  if ((e = hipStreamSynchronize(m->stream))) {
    delete m;
    return hip_err(e);
  }
  *out = m;
  return NPHYS_OK;
}

int nphys_mlp_destroy(nphys_mlp* m) {
  if (!m) return set_err(NPHYS_ERR_INVALID, "null network");
  delete m;
  return NPHYS_OK;
}

int nphys_mlp_num_params(int64_t* n) {
  if (!n) return set_err(NPHYS_ERR_INVALID, "null argument");
  *n = kNumParams;
  return NPHYS_OK;
}

int nphys_mlp_get_params(nphys_mlp* m, float* out, int64_t cap) {
  if (!m || !out) return set_err(NPHYS_ERR_INVALID, "null argument");
  if (cap < kNumParams) return set_err(NPHYS_ERR_CAPACITY, "parameter buffer too small");
  HIPCK(hipMemcpyAsync(out, m->params.p, kNumParams * 4, hipMemcpyDeviceToHost, m->stream));
  HIPCK(hipStreamSynchronize(m->stream));
  return NPHYS_OK;
}

int nphys_mlp_set_params(nphys_mlp* m, const float* in, int64_t n) {
  if (!m || !in) return set_err(NPHYS_ERR_INVALID, "null argument");
  if (n != kNumParams) return set_err(NPHYS_ERR_INVALID, "parameter count mismatch");
  for (int i = 0; i < kFeat; ++i)
    if (!(in[kOffXScale + i] > 0.0f)) return set_err(NPHYS_ERR_INVALID, "feature scales must be > 0");
  for (int i = 0; i < kOut; ++i)
    if (!(in[kOffYScale + i] > 0.0f)) return set_err(NPHYS_ERR_INVALID, "output scales must be > 0");
  HIPCK(hipMemcpyAsync(m->params.p, in, kNumParams * 4, hipMemcpyHostToDevice, m->stream));
  HIPCK(hipMemsetAsync(m->mom1.p, 0, kNumTrained * 4, m->stream));
  HIPCK(hipMemsetAsync(m->mom2.p, 0, kNumTrained * 4, m->stream));
  mlp_pack_images(m->params.as<float>(), m->img.as<uint16_t>(), m->stream);
  HIPCK(hipStreamSynchronize(m->stream));
  m->adam_t = 0;
  return NPHYS_OK;
}

int nphys_mlp_set_data(nphys_mlp* m, const float* X, const float* T, int64_t n, int32_t set_normalisation) {
  if (!m || (!X && n) || (!T && n) || n < 0) return set_err(NPHYS_ERR_INVALID, "bad dataset");
  if (n > (int64_t(1) << 30)) return set_err(NPHYS_ERR_CAPACITY, "dataset too large");
  if (set_normalisation && n > 0) {
    // per-feature mean and standard deviation, per-output root mean square (double sums)
    std::vector<float> p(kNumParams);
    HIPCK(hipMemcpy(p.data(), m->params.p, kNumParams * 4, hipMemcpyDeviceToHost));
    for (int f = 0; f < kFeat; ++f) {
      double s = 0.0, ss = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        const double v = X[i * kFeat + f];
        s += v;
        ss += v * v;
      }
      const double mean = s / double(n);
      const double var = std::max(0.0, ss / double(n) - mean * mean);
      p[kOffXMean + f] = float(mean);
      p[kOffXScale + f] = float(var > 1e-12 ? std::sqrt(var) : 1.0);
    }
    for (int j = 0; j < kOut; ++j) {
      double ss = 0.0;
      for (int64_t i = 0; i < n; ++i) ss += double(T[i * kOut + j]) * double(T[i * kOut + j]);
      const double rms = std::sqrt(ss / double(n));
      p[kOffYScale + j] = float(rms > 1e-12 ? rms : 1.0);
    }
    HIPCK(hipMemcpy(m->params.p, p.data(), kNumParams * 4, hipMemcpyHostToDevice));
  }
  const size_t nn = size_t(std::max<int64_t>(n, 1));
  HIPCK(m->X.ensure(nn * kFeat * 4));
  HIPCK(m->T.ensure(nn * kOut * 4));
  HIPCK(m->perm.ensure(nn * 4));
  if (n > 0) {
    HIPCK(hipMemcpy(m->X.p, X, size_t(n) * kFeat * 4, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(m->T.p, T, size_t(n) * kOut * 4, hipMemcpyHostToDevice));
  }
  m->n = n;
  return NPHYS_OK;
}

int nphys_mlp_train(nphys_mlp* m, int32_t epochs, int32_t batch, float lr, float* epoch_loss) {
  if (!m) return set_err(NPHYS_ERR_INVALID, "null network");
  if (epochs < 0 || batch < 64 || batch % 64 || batch > (1 << 16) || !(lr > 0.0f))
    return set_err(NPHYS_ERR_INVALID, "epochs >= 0, batch a multiple of 64 in [64, 65536], lr > 0");
  if (m->n < batch) return set_err(NPHYS_ERR_STATE, "dataset smaller than one batch");
  HIPCK(m->ensure_batch(batch));
  const long nb = m->n / batch;
  HIPCK(m->losses.ensure(size_t(nb) * 8));
  std::vector<int> perm(static_cast<size_t>(m->n));
  std::vector<double> losses(static_cast<size_t>(nb));
  for (int ep = 0; ep < epochs; ++ep) {
    // Fisher-Yates over SplitMix64 seeded by (seed, epoch)
    uint64_t s = m->seed * 0x100000001B3ull + uint64_t(ep) + 1;
    for (long i = 0; i < m->n; ++i) perm[size_t(i)] = int(i);
    for (long i = m->n - 1; i > 0; --i) std::swap(perm[size_t(i)], perm[size_t(splitmix(s) % uint64_t(i + 1))]);
    HIPCK(hipMemcpyAsync(m->perm.p, perm.data(), size_t(m->n) * 4, hipMemcpyHostToDevice, m->stream));
    for (long k = 0; k < nb; ++k) {
      MlpBatch b = m->batch(batch, m->X.as<float>(), m->T.as<float>(), m->perm.as<int>() + k * batch);
      mlp_batch_gradients(m->dev(), b, m->grad.as<float>(), m->losses.as<double>() + k, m->stream);
      ++m->adam_t;
      mlp_adam(m->params.as<float>(), m->mom1.as<float>(), m->mom2.as<float>(), m->grad.as<float>(), lr, 0.9f,
               0.999f, 1e-8f, int(std::min<long>(m->adam_t, 1 << 30)), m->img.as<uint16_t>(), m->stream);
    }
    HIPCK(hipMemcpyAsync(losses.data(), m->losses.p, size_t(nb) * 8, hipMemcpyDeviceToHost, m->stream));
    HIPCK(hipStreamSynchronize(m->stream));
    if (epoch_loss) {
      double l = 0.0;
      for (double v : losses) l += v;
      epoch_loss[ep] = float(l / double(nb));
    }
  }
  HIPCK(hipStreamSynchronize(m->stream));
  return NPHYS_OK;
}

int nphys_mlp_forward(nphys_mlp* m, const float* X, int64_t n, float* out) {
  if (!m || ((!X || !out) && n) || n < 0) return set_err(NPHYS_ERR_INVALID, "bad argument");
  if (n == 0) return NPHYS_OK;
  if (n > (int64_t(1) << 30)) return set_err(NPHYS_ERR_CAPACITY, "too many rows");
  HIPCK(m->fx.ensure(size_t(n) * kFeat * 4));
  HIPCK(m->fy.ensure(size_t(n) * kOut * 4));
  HIPCK(hipMemcpyAsync(m->fx.p, X, size_t(n) * kFeat * 4, hipMemcpyHostToDevice, m->stream));
  mlp_infer(m->dev(), m->fx.as<float>(), int(n), m->fy.as<float>(), m->stream);
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(out, m->fy.p, size_t(n) * kOut * 4, hipMemcpyDeviceToHost, m->stream));
  HIPCK(hipStreamSynchronize(m->stream));
  return NPHYS_OK;
}

int nphys_mlp_gradients(nphys_mlp* m, const float* X, const float* T, int64_t n, float* grad, double* loss) {
  if (!m || !X || !T || !grad || !loss) return set_err(NPHYS_ERR_INVALID, "null argument");
  if (n < 1 || n > (1 << 16)) return set_err(NPHYS_ERR_INVALID, "rows must be in [1, 65536]");
  const int B = int(n);
  HIPCK(m->ensure_batch(((B + 63) / 64) * 64));
  HIPCK(m->fx.ensure(size_t(n) * kFeat * 4));
  HIPCK(m->fy.ensure(std::max(size_t(n) * kOut * 4, size_t(n) * 4)));
  DBuf tt, id;
  HIPCK(tt.ensure(size_t(n) * kOut * 4));
  HIPCK(id.ensure(size_t(n) * 4));
  std::vector<int> iota(static_cast<size_t>(n));
  for (int i = 0; i < B; ++i) iota[size_t(i)] = i;
  DBuf dl;
  HIPCK(dl.ensure(8));
  HIPCK(hipMemcpyAsync(m->fx.p, X, size_t(n) * kFeat * 4, hipMemcpyHostToDevice, m->stream));
  HIPCK(hipMemcpyAsync(tt.p, T, size_t(n) * kOut * 4, hipMemcpyHostToDevice, m->stream));
  HIPCK(hipMemcpyAsync(id.p, iota.data(), size_t(n) * 4, hipMemcpyHostToDevice, m->stream));
  MlpBatch b = m->batch(B, m->fx.as<float>(), tt.as<float>(), id.as<int>());
  mlp_batch_gradients(m->dev(), b, m->grad.as<float>(), dl.as<double>(), m->stream);
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(grad, m->grad.p, kNumTrained * 4, hipMemcpyDeviceToHost, m->stream));
  HIPCK(hipMemcpyAsync(loss, dl.p, 8, hipMemcpyDeviceToHost, m->stream));
  HIPCK(hipStreamSynchronize(m->stream));
  return NPHYS_OK;
}

int nphys_mlp_gradients_noalloc(nphys_mlp* m, const float* X, const float* T, int64_t n, float* grad, double* loss,
                                void* stream) {
  if (!m || !X || !T || !grad || !loss) return set_err(NPHYS_ERR_INVALID, "null argument");
  if (n < 1 || n > kNoAllocMaxRows)
    return set_err(NPHYS_ERR_CAPACITY, "rows must be in [1, " + std::to_string(kNoAllocMaxRows) +
                                           "] for nphys_mlp_gradients_noalloc");
  unsigned char* base = nullptr;
  if (const int rc = scratch_on_current_device(&base)) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : m->stream;
  MlpBatch b;
  b.B = int(n);
  b.X = X;
  b.T = T;
  b.idx = reinterpret_cast<const int*>(base + kScrIdx);
  b.A0 = reinterpret_cast<uint16_t*>(base + kScrA0);
  b.Z1 = reinterpret_cast<float*>(base + kScrZ1);
  b.Z2 = reinterpret_cast<float*>(base + kScrZ2);
  b.H1 = reinterpret_cast<uint16_t*>(base + kScrH1);
  b.H2 = reinterpret_cast<uint16_t*>(base + kScrH2);
  b.Y = reinterpret_cast<float*>(base + kScrY);
  b.D3 = reinterpret_cast<uint16_t*>(base + kScrD3);
  b.D2 = reinterpret_cast<uint16_t*>(base + kScrD2);
  b.D1 = reinterpret_cast<uint16_t*>(base + kScrD1);
  b.loss_part = reinterpret_cast<double*>(base + kScrLoss);
  b.gpart = reinterpret_cast<float*>(base + kScrGpart);
  hipLaunchKernelGGL(k_mlp_iota, dim3((b.B + 255) / 256), dim3(256), 0, s, reinterpret_cast<int*>(base + kScrIdx),
                     b.B);
  mlp_batch_gradients(m->dev(), b, grad, loss, s);
  HIPCK(hipGetLastError());
  if (!stream) HIPCK(hipStreamSynchronize(s));
  return NPHYS_OK;
}

}  // extern "C"

// nphysics_amd/csrc/k_util.hip -- device-wide exclusive scan and stable LDS radix sort.
//
// Scan: 256-thread blocks each own a tile of 2048 words (8 per thread); block totals are scanned
// recursively and added back. Radix sort: 8-bit digits; per pass a histogram kernel (LDS
// atomics, one 256-bin row per tile, digit-major so that one exclusive scan yields every tile's
// output base) and a scatter kernel that ranks keys stably inside a tile with 64-lane wavefront
// ballots (eight ballots give each lane the set of lanes holding the same digit).
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace nph {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

constexpr int kRadixThreads = 256;
constexpr int kRadixItems = 8;
constexpr int kRadixTile = kRadixThreads * kRadixItems;
constexpr int kRadixBins = 256;

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const uint32_t* __restrict__ in,
                                                            uint32_t* __restrict__ out, int n,
                                                            uint32_t* __restrict__ tile_sums) {
  __shared__ uint32_t s_part[kScanThreads];
  const int tid = threadIdx.x;
  const long base = long(blockIdx.x) * kScanTile + long(tid) * kScanItems;
  uint32_t v[kScanItems];
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const long i = base + k;
    v[k] = i < n ? in[i] : 0u;
    acc += v[k];
  }
  s_part[tid] = acc;
  __syncthreads();
  // Hillis-Steele inclusive scan of the per-thread sums.
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const uint32_t x = tid >= off ? s_part[tid - off] : 0u;
    __syncthreads();
    s_part[tid] += x;
    __syncthreads();
  }
  uint32_t run = tid > 0 ? s_part[tid - 1] : 0u;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const long i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (tid == kScanThreads - 1 && tile_sums) tile_sums[blockIdx.x] = s_part[tid];
}

__global__ void k_scan_add(uint32_t* __restrict__ out, int n, const uint32_t* __restrict__ tile_offs) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] += tile_offs[i / kScanTile];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(kRadixThreads) k_radix_hist(const uint32_t* __restrict__ keys,
                                                             int n, int shift, int ntiles,
                                                             uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_h[kRadixBins];
  const int tid = threadIdx.x;
  s_h[tid] = 0;
  __syncthreads();
  const long base = long(blockIdx.x) * kRadixTile;
#pragma unroll
  for (int it = 0; it < kRadixItems; ++it) {
    const long i = base + it * kRadixThreads + tid;
    if (i < n) atomicAdd(&s_h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[long(tid) * ntiles + blockIdx.x] = s_h[tid];
}

__global__ void __launch_bounds__(kRadixThreads) k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int n, int shift, int ntiles, const uint32_t* __restrict__ offs) {
  __shared__ uint32_t s_base[kRadixBins];
  __shared__ uint32_t s_wcnt[kRadixThreads / 64][kRadixBins];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  s_base[tid] = offs[long(tid) * ntiles + blockIdx.x];
  const long base = long(blockIdx.x) * kRadixTile;
  for (int it = 0; it < kRadixItems; ++it) {
#pragma unroll
    for (int w = 0; w < kRadixThreads / 64; ++w) s_wcnt[w][tid] = 0;
    __syncthreads();
    const long i = base + it * kRadixThreads + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const uint32_t lrank = __popcll(m & lt_mask);
    if (valid && lrank == 0) s_wcnt[wave][d] = uint32_t(__popcll(m));
    __syncthreads();
    if (valid) {
      uint32_t pre = 0;
      for (int w = 0; w < wave; ++w) pre += s_wcnt[w][d];
      const uint32_t pos = s_base[d] + pre + lrank;
      kout[pos] = key;
      vout[pos] = vin[i];
    }
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kRadixThreads / 64; ++w) tot += s_wcnt[w][tid];
    s_base[tid] += tot;
    __syncthreads();
  }
}

int div_up(long a, long b) { return int((a + b - 1) / b); }

}  // namespace

size_t scan_temp_words(int n) {
  size_t words = 0;
  long m = n;
  while (m > kScanTile) {
    const long t = (m + kScanTile - 1) / kScanTile;
    words += 2 * size_t(t);
    m = t;
  }
  return words + 16;
}

void scan_exclusive_u32(const uint32_t* in, uint32_t* out, int n, uint32_t* tmp, hipStream_t s) {
  if (n <= 0) return;
  const int tiles = div_up(n, kScanTile);
  if (tiles == 1) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, s, in, out, n, nullptr);
    return;
  }
  uint32_t* sums = tmp;
  uint32_t* offs = tmp + tiles;
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanThreads), 0, s, in, out, n, sums);
  scan_exclusive_u32(sums, offs, tiles, tmp + 2 * size_t(tiles), s);
  hipLaunchKernelGGL(k_scan_add, dim3(div_up(n, 256)), dim3(256), 0, s, out, n, offs);
}

size_t radix_temp_words(int n) {
  const int ntiles = div_up(n > 0 ? n : 1, kRadixTile);
  const size_t h = size_t(kRadixBins) * ntiles;
  return 2 * h + scan_temp_words(int(h)) + 16;
}

void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                      int n, int bits, uint32_t* tmp, hipStream_t s) {
  if (n <= 1) return;
  const int ntiles = div_up(n, kRadixTile);
  const size_t h = size_t(kRadixBins) * ntiles;
  uint32_t* hist = tmp;
  uint32_t* offs = tmp + h;
  uint32_t* stmp = tmp + 2 * h;
  uint32_t* kin = keys;
  uint32_t* vin = vals;
  uint32_t* kout = keys_alt;
  uint32_t* vout = vals_alt;
  int passes = 0;
  for (int shift = 0; shift < bits; shift += 8, ++passes) {
    hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kRadixThreads), 0, s, kin, n, shift, ntiles, hist);
    scan_exclusive_u32(hist, offs, int(h), stmp, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(kRadixThreads), 0, s, kin, vin, kout, vout, n,
                       shift, ntiles, offs);
    uint32_t* t = kin; kin = kout; kout = t;
    t = vin; vin = vout; vout = t;
  }
  if (passes & 1) {
    (void)hipMemcpyAsync(keys, kin, size_t(n) * 4, hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(vals, vin, size_t(n) * 4, hipMemcpyDeviceToDevice, s);
  }
}

void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, p, v, n);
}

}  // namespace nph

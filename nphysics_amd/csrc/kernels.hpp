// nphysics_amd/csrc/kernels.hpp -- launch wrappers of the HIP kernels (host side).
//
// Every wrapper enqueues on the given stream and returns immediately. Device pointers only.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace nph {

// ---------------------------------------------------------------- utilities (k_util.hip)
// Exclusive prefix sum of n uint32 values; `tmp` must hold scan_temp_words(n) words.
size_t scan_temp_words(int n);
void scan_exclusive_u32(const uint32_t* in, uint32_t* out, int n, uint32_t* tmp, hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the key.
// Result ends in (keys, vals); (keys_alt, vals_alt) are scratch of the same size.
// `tmp` must hold radix_temp_words(n) words.
size_t radix_temp_words(int n);
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                      int n, int bits, uint32_t* tmp, hipStream_t s);

void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

}  // namespace nph

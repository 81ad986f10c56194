// Device-wide primitives: exclusive scan and a stable LSD radix sort.
#pragma once
#include "common.hpp"

namespace pyfof {

// Scratch sizes (bytes) the primitives need for n elements.
size_t scan_temp_bytes(uint64_t n);
size_t radix_temp_bytes(uint64_t n);

// out[i] = sum(in[0..i)), i < n; if total != nullptr, *total (device) = sum(in)
// (n must be > 0 for total to be written).
// in and out may alias. temp: scan_temp_bytes(n) bytes of device memory.
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total,
                        void* temp, hipStream_t s);

// out[i] = max(in[0..i]) (inclusive), i < n. in and out may alias.
void inclusive_max_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* temp,
                            hipStream_t s);

// Stable ascending sort of (key, value) pairs on the low `bits` key bits.
// keys/vals hold the input; alt_keys/alt_vals are same-size buffers. Returns
// true if the sorted result ended in the alt buffers (odd number of passes).
// With vals_iota the input values are taken to be 0, 1, .., n - 1 without
// reading vals (whose contents are then ignored; it still serves as a
// buffer).
bool radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* alt_keys, uint32_t* alt_vals,
                      uint64_t n, int bits, void* temp, hipStream_t s, bool vals_iota = false);

// bit width needed to represent values in [0, max_key]
int bits_for(uint64_t max_key);

}  // namespace pyfof

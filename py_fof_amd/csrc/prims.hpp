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
//
// With first_hist_ready the first pass's tile histogram is already in
// radix_first_hist(temp) (written by the caller, as the key producer can
// count its keys while it writes them): digit d of tile t (keys
// [t * radix_tile_keys, (t + 1) * radix_tile_keys)) at entry d * ntiles + t,
// ntiles = ceil(n / radix_tile_keys), the digit being the key's low
// radix_first_width(bits) bits.
bool radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* alt_keys, uint32_t* alt_vals,
                      uint64_t n, int bits, void* temp, hipStream_t s, bool vals_iota = false,
                      bool first_hist_ready = false);

// keys per tile of the radix passes, the first pass's digit width for a
// sort on `bits` key bits (0: no pass), and where in temp its histogram goes
constexpr uint32_t radix_tile_keys = 8192;
int radix_first_width(int bits);
uint32_t* radix_first_hist(void* temp);

// bit width needed to represent values in [0, max_key]
int bits_for(uint64_t max_key);

}  // namespace pyfof

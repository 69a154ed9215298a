// radix_sort.hpp -- stable LSD radix sort of 64-bit keys (optional 32-bit values)
// on gfx950: 8-bit digits; per pass a tile histogram, a per-digit scan over tiles
// and a scatter that ranks keys inside each wave with ballots (wave64
// multisplit), so equal digits keep their order (stable).
#pragma once
#include "engine.hpp"

namespace dcg {

// Device scratch (bytes) a sort of n keys needs (with values when with_values).
size_t radix_sort_temp_bytes(uint64_t n, bool with_values);

// Sort keys_in[0..n) ascending by bits [begin_bit, end_bit) into keys_out
// (values carried from vals_in to vals_out when both are given).  keys_in and
// vals_in are left unchanged.  Enqueued on `s`.
void radix_sort(hipStream_t s, const uint64_t *keys_in, uint64_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out,
                uint64_t n, int begin_bit, int end_bit, void *temp, size_t temp_bytes);

}  // namespace dcg

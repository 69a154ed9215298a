// primitives.hpp -- device-wide sort, scan and selection, hand-written for
// gfx950 (wave64).  Every routine is enqueued on a caller's stream and works
// in caller scratch (the *_temp_bytes functions size it), so none allocates.
#pragma once
#include "engine.hpp"

namespace dcg {

// ---- stable LSD radix sort ------------------------------------------------
// Keys K = uint32_t or uint64_t, optional uint32_t values.  Sorts
// keys_in[0..n) ascending by bits [begin_bit, end_bit) into keys_out (values
// from vals_in to vals_out when both are given); equal digits keep their input
// order, so the sort is stable and a range that leaves bits out orders by the
// range only.  keys_in / vals_in are left unchanged; in and out must differ.
template <typename K>
size_t radix_sort_temp_bytes(uint64_t n, bool with_values);
template <typename K>
void radix_sort(hipStream_t s, const K *keys_in, K *keys_out, const uint32_t *vals_in, uint32_t *vals_out, uint64_t n,
                int begin_bit, int end_bit, void *temp, size_t temp_bytes);

// Keys whose ascending order is the descending order of the floats in[]
// (-0 and +0 equal): for a stable descending sort of float scores.
void float_desc_keys(hipStream_t s, const float *in, uint32_t *out, uint64_t n);

// ---- exclusive prefix sum -------------------------------------------------
// out[i] = in[0] + ... + in[i-1] for T = uint32_t or uint64_t (wrapping), in
// and out distinct.
template <typename T>
size_t scan_temp_bytes(uint64_t n);
template <typename T>
void exclusive_sum(hipStream_t s, const T *in, T *out, uint64_t n, void *temp, size_t temp_bytes);

// ---- selection ------------------------------------------------------------
// The indices i with flags[i] != 0, in ascending order, into out; their
// number into *count (device memory).
size_t select_temp_bytes(uint64_t n);
void select_flagged(hipStream_t s, const uint8_t *flags, uint64_t n, uint32_t *out, unsigned long long *count,
                    void *temp, size_t temp_bytes);

}  // namespace dcg

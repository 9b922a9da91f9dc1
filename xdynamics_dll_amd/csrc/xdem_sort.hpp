// xdem_sort.hpp -- stable LSD radix sort of (uint32 key, uint32 value) pairs on gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace xdem {

// Bytes of scratch the sort needs for n pairs (histograms only; the caller provides
// the key/value ping-pong buffers).
size_t radix_sort_scratch_bytes(uint32_t n);

// Sorts n pairs by the low `bits` bits of the key (1..32), stable: equal keys keep
// their input order. Reads keys_in/vals_in (unchanged), writes the result to
// keys_out/vals_out, and uses keys_tmp/vals_tmp and `scratch` as work space. All
// pointers are device pointers; the work is queued on `stream`.
hipError_t radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                            uint32_t* vals_out, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
                            int bits, void* scratch, hipStream_t stream);

}  // namespace xdem

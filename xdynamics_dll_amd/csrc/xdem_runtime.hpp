// xdem_runtime.hpp -- process-wide state of the launcher-level C ABI: the parameter set
// (the reference keeps it in a __constant__ symbol, def:16-19), quirk switches, the stream,
// the wall/mesh body-force buffers (xContact's static buffers, xContact.cpp:339-345),
// and per-device scratch (sort buffers, reduction partials, the device error word).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "xdem_types.h"

namespace xdem {

struct Runtime {
    xdem_dem_parameters prm;
    bool prm_set = false;
    xdem_quirks quirks = {0, 0, 1, 1};
    hipStream_t stream = nullptr;
    double* bf[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // scratch owned by the library (grown on demand; launchers may allocate, as the
    // reference's thrust calls do)
    uint32_t* err = nullptr;          // device error word
    void* sort_scratch = nullptr;
    size_t sort_scratch_bytes = 0;
    uint32_t* sort_buf = nullptr;     // 4 arrays of sort_cap entries
    size_t sort_cap = 0;
    double* red_partial = nullptr;    // reduction partials
    double* own_bf = nullptr;         // body-force buffers used when none are registered
    size_t own_bf_n = 0;
};

Runtime& runtime();
// Makes sure the runtime scratch holds at least n sort entries and nbf body-force entries.
hipError_t runtime_reserve(size_t n_sort, size_t n_bf);

#define XDEM_HIP_CHECK(expr)                         \
    do {                                             \
        hipError_t e_ = (expr);                      \
        if (e_ != hipSuccess) return XDEM_E_HIP;     \
    } while (0)

}  // namespace xdem

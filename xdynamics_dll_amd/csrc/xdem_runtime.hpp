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
    // Workspace of the launchers (device error word, reduction partials, sort buffers and
    // histograms, body-force buffers used when none are registered) for launches of up to
    // ws_n elements: reserved by the library at xdem_setDEMSymbolicParameter or
    // xdem_reserve_workspace (init calls), or owned by the caller (xdem_set_workspace). No
    // launcher allocates: one that needs more than the workspace holds returns an error.
    char* ws = nullptr;
    size_t ws_bytes = 0;
    uint32_t ws_n = 0;
    bool ws_owned = false;            // allocated by the library (freed when replaced)
    uint32_t* err = nullptr;          // device error word
    double* red_partial = nullptr;    // reduction partials (1024)
    uint32_t* sort_buf = nullptr;     // 4 arrays of ws_n entries
    void* sort_scratch = nullptr;
    double* own_bf = nullptr;         // 6 arrays of ws_n entries
    unsigned host_err = 0;            // XDEM_ERR_WORKSPACE, reported with the device word
};

Runtime& runtime();

// Bytes of the launcher workspace for launches of up to n elements.
size_t workspace_bytes(uint32_t n);

#define XDEM_HIP_CHECK(expr)                         \
    do {                                             \
        hipError_t e_ = (expr);                      \
        if (e_ != hipSuccess) return XDEM_E_HIP;     \
    } while (0)

}  // namespace xdem

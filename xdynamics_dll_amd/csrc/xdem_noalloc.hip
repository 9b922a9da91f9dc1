// xdem_noalloc.hip -- xdem_cu_reorderDataAndFindCellStart_noalloc: the cell sort of the
// launcher layer for callers that cannot allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation). Its sort buffers and digit histograms
// live in a file-scope device array instead of the launcher workspace; it calls no
// allocating routine. Same results as xdem_cu_reorderDataAndFindCellStart.
#include <mutex>

#include "xdem_capi.h"
#include "xdem_runtime.hpp"
#include "xdem_sort.hpp"

namespace xdem {
namespace {

constexpr uint32_t kNoAllocMaxN = 1u << 20;                              // elements per call
constexpr size_t kSortWords = 4ull * kNoAllocMaxN;                       // keys and values, two buffers each
constexpr size_t kHistWords = 256ull * ((kNoAllocMaxN + 4095) / 4096);   // per-tile digit counts
constexpr size_t kScratchWords = kSortWords + kHistWords;                // 16.25 MiB

// The scratch: one copy per GPU, held until the process ends. Host code never passes this
// name to a kernel or a copy: it uses the device address looked up below.
static __device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_ptr[kMaxDevices] = {};

// Device address of the current GPU's scratch, looked up on the first call on that GPU.
int scratch_address(uint32_t** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return XDEM_E_HIP;
    if (dev < 0 || dev >= kMaxDevices) return XDEM_E_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return XDEM_E_HIP;
        g_ptr[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_ptr[dev];
    return XDEM_OK;
}

constexpr unsigned kThreads = 256;

__global__ __launch_bounds__(kThreads) void k_clear_cells(unsigned* cstart, unsigned* cend, size_t ncell)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncell) {
        cstart[i] = 0xffffffffu;
        cend[i] = 0u;
    }
}

// First and one-past-last sorted position of every occupied cell (impl:317-360).
__global__ __launch_bounds__(kThreads) void k_bounds(const unsigned* hash, const unsigned* index, unsigned* cstart,
                                                     unsigned* cend, unsigned* sorted_index, unsigned n)
{
    const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const unsigned h = hash[id];
    if (id == 0 || h != hash[id - 1]) {
        cstart[h] = id;
        if (id > 0) cend[hash[id - 1]] = id;
    }
    if (id == n - 1) cend[h] = id + 1;
    sorted_index[id] = index[id];
}

}  // namespace
}  // namespace xdem

using namespace xdem;

extern "C" int xdem_cu_reorderDataAndFindCellStart_noalloc(unsigned int* hash, unsigned int* index,
                                                           unsigned int* cstart, unsigned int* cend,
                                                           unsigned int* sorted_index, unsigned int np,
                                                           unsigned int ncell)
{
    if (!hash || !index || !cstart || !cend || !sorted_index || ncell == 0) return XDEM_E_INVALID;
    if (np > kNoAllocMaxN) return XDEM_E_INVALID;   // beyond the scratch: refused, nothing queued
    uint32_t* s = nullptr;
    const int rc = scratch_address(&s);
    if (rc) return rc;
    hipStream_t st = runtime().stream;
    if (np > 0) {
        int bits = 1;
        while (bits < 32 && (1ull << bits) < (unsigned long long)ncell) bits++;
        uint32_t* kout = s;
        uint32_t* vout = s + kNoAllocMaxN;
        uint32_t* ktmp = s + 2 * (size_t)kNoAllocMaxN;
        uint32_t* vtmp = s + 3 * (size_t)kNoAllocMaxN;
        uint32_t* hist = s + kSortWords;   // fully rewritten by every pass: no clearing needed
        XDEM_HIP_CHECK(radix_sort_pairs(hash, index, kout, vout, ktmp, vtmp, np, bits, hist, st));
        XDEM_HIP_CHECK(hipMemcpyAsync(hash, kout, sizeof(unsigned) * np, hipMemcpyDeviceToDevice, st));
        XDEM_HIP_CHECK(hipMemcpyAsync(index, vout, sizeof(unsigned) * np, hipMemcpyDeviceToDevice, st));
    }
    hipLaunchKernelGGL(k_clear_cells, dim3((unsigned)((ncell + kThreads - 1) / kThreads)), dim3(kThreads), 0, st,
                       cstart, cend, (size_t)ncell);
    if (np > 0)
        hipLaunchKernelGGL(k_bounds, dim3((np + kThreads - 1) / kThreads), dim3(kThreads), 0, st, hash, index, cstart,
                           cend, sorted_index, np);
    if (hipGetLastError() != hipSuccess) return XDEM_E_HIP;
    return XDEM_OK;
}

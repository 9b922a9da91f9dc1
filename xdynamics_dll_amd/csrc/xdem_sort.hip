// xdem_sort.hip -- stable LSD radix sort of (uint32 key, uint32 value) pairs for gfx950.
//
// 8-bit digits, three kernels per digit:
//   k_rs_hist    per-tile digit counts (LDS atomics) -> hist[digit][tile]
//   k_rs_scan    one work-group exclusive scan of hist (digit-major, so the result is the
//                global output offset of each (digit, tile))
//   k_rs_scatter ranks every key of a tile among equal digits with wave64 ballots (the
//                rank respects input order, which makes the sort stable), reorders the
//                tile in LDS by digit, and writes runs of equal digits to global memory.
// A tile is 256 threads x 16 keys = 4096 keys.
#include "xdem_sort.hpp"

namespace xdem {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;
constexpr int kWaves = kBlock / 64;

__global__ __launch_bounds__(kBlock) void k_rs_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                   uint32_t* __restrict__ hist, uint32_t ntiles)
{
    __shared__ uint32_t cnt[256];
    const uint32_t tid = threadIdx.x;
    cnt[tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile;
#pragma unroll
    for (int s = 0; s < kItems; s++) {
        const uint32_t idx = base + s * kBlock + tid;
        if (idx < n) atomicAdd(&cnt[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[tid * ntiles + blockIdx.x] = cnt[tid];
}

// Exclusive scan of `total` entries (a multiple of 4) by one work-group of 1024 threads.
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t* __restrict__ data, uint32_t total)
{
    __shared__ uint32_t s_wsum[16];
    __shared__ uint32_t s_carry;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < total; base += 4096u) {
        const uint32_t idx = base + tid * 4u;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (idx < total) v = *reinterpret_cast<const uint4*>(data + idx);
        const uint32_t t0 = v.x, t1 = t0 + v.y, t2 = t1 + v.z, t3 = t2 + v.w;
        uint32_t x = t3;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= (uint32_t)off) x += y;
        }
        if (lane == 63) s_wsum[w] = x;
        __syncthreads();
        if (w == 0) {
            uint32_t s = lane < 16 ? s_wsum[lane] : 0u;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                const uint32_t y = __shfl_up(s, off, 64);
                if (lane >= (uint32_t)off) s += y;
            }
            if (lane < 16) s_wsum[lane] = s;
        }
        __syncthreads();
        const uint32_t pre = s_carry + (w > 0 ? s_wsum[w - 1] : 0u) + (x - t3);
        if (idx < total) *reinterpret_cast<uint4*>(data + idx) = make_uint4(pre, pre + t0, pre + t1, pre + t2);
        __syncthreads();
        if (tid == 0) s_carry += s_wsum[15];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                      uint32_t n, int shift, const uint32_t* __restrict__ offsets,
                                                      uint32_t ntiles)
{
    __shared__ uint32_t s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];
    __shared__ uint32_t s_cnt[2][kWaves][256];
    __shared__ uint32_t s_run[256];
    __shared__ uint32_t s_start[256];
    __shared__ uint32_t s_gbase[256];

    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t base = blockIdx.x * kTile;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));

    s_run[tid] = 0;
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < kWaves; q++) s_cnt[b][q][tid] = 0;
    s_gbase[tid] = offsets[tid * ntiles + blockIdx.x];

    uint32_t key[kItems], val[kItems], rank[kItems];
#pragma unroll
    for (int s = 0; s < kItems; s++) {
        const uint32_t idx = base + s * kBlock + tid;
        key[s] = idx < n ? kin[idx] : 0xffffffffu;
        val[s] = idx < n ? vin[idx] : 0u;
    }
    __syncthreads();

#pragma unroll
    for (int s = 0; s < kItems; s++) {
        const uint32_t buf = s & 1;
        const uint32_t idx = base + s * kBlock + tid;
        const uint32_t d = idx < n ? ((key[s] >> shift) & 255u) : 255u;
        // lanes of this wave holding the same digit
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t before = (uint32_t)__popcll(peers & lt_mask);
        if (before == 0) s_cnt[buf][w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            // thread `tid` owns digit `tid`: wave counts -> exclusive bases; clear the other buffer
            uint32_t run = s_run[tid];
#pragma unroll
            for (int q = 0; q < kWaves; q++) {
                const uint32_t c = s_cnt[buf][q][tid];
                s_cnt[buf][q][tid] = run;
                run += c;
            }
            s_run[tid] = run;
#pragma unroll
            for (int q = 0; q < kWaves; q++) s_cnt[buf ^ 1u][q][tid] = 0;
        }
        __syncthreads();
        rank[s] = s_cnt[buf][w][d] + before;
    }
    __syncthreads();

    // exclusive scan of the tile's digit counts (s_run) -> s_start
    {
        const uint32_t c = s_run[tid];
        uint32_t x = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= (uint32_t)off) x += y;
        }
        __shared__ uint32_t s_wtot[kWaves];
        if (lane == 63) s_wtot[w] = x;
        __syncthreads();
        uint32_t pre = 0;
        for (uint32_t q = 0; q < w; q++) pre += s_wtot[q];
        s_start[tid] = pre + x - c;
    }
    __syncthreads();

#pragma unroll
    for (int s = 0; s < kItems; s++) {
        const uint32_t idx = base + s * kBlock + tid;
        const uint32_t d = idx < n ? ((key[s] >> shift) & 255u) : 255u;
        const uint32_t p = s_start[d] + rank[s];
        s_keys[p] = key[s];
        s_vals[p] = val[s];
    }
    __syncthreads();

    const uint32_t valid = n - base < (uint32_t)kTile ? n - base : (uint32_t)kTile;
#pragma unroll
    for (int s = 0; s < kItems; s++) {
        const uint32_t p = s * kBlock + tid;
        if (p < valid) {
            const uint32_t k = s_keys[p];
            const uint32_t d = (k >> shift) & 255u;
            const uint32_t o = s_gbase[d] + (p - s_start[d]);
            kout[o] = k;
            vout[o] = s_vals[p];
        }
    }
}

}  // namespace

size_t radix_sort_scratch_bytes(uint32_t n)
{
    const size_t ntiles = (n + kTile - 1) / kTile;
    return sizeof(uint32_t) * 256 * (ntiles ? ntiles : 1);
}

hipError_t radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                            uint32_t* vals_out, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
                            int bits, void* scratch, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    if (bits < 1) bits = 1;
    if (bits > 32) bits = 32;
    const int npass = (bits + 7) / 8;
    const uint32_t ntiles = (n + kTile - 1) / kTile;
    uint32_t* hist = static_cast<uint32_t*>(scratch);
    const uint32_t* ksrc = keys_in;
    const uint32_t* vsrc = vals_in;
    for (int p = 0; p < npass; p++) {
        // the last pass must land in *_out
        const bool to_out = ((npass - 1 - p) % 2) == 0;
        uint32_t* kdst = to_out ? keys_out : keys_tmp;
        uint32_t* vdst = to_out ? vals_out : vals_tmp;
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kBlock), 0, stream, ksrc, n, shift, hist, ntiles);
        hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(1024), 0, stream, hist, 256u * ntiles);
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kBlock), 0, stream, ksrc, vsrc, kdst, vdst, n, shift,
                           hist, ntiles);
        ksrc = kdst;
        vsrc = vdst;
    }
    return hipGetLastError();
}

}  // namespace xdem

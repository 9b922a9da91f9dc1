// xdem_grid.hip -- the bucket grid and neighbour lists of the fused step (xdem_ctx path).
//
// The reference sorts (hash, index) pairs with a radix sort every step and then finds the
// first and last sorted position of every cell (def:94-124; impl:317-360). Here the grid is
// a counting sort over a dense bucket table, rebuilt together with the neighbour lists only
// on the steps that need it: the lists hold every pair within reach plus a skin, and the
// drift of each step (drift_position) flags the next step for a rebuild once a sphere has
// moved half the skin since the last build. Every kernel below returns at once on the other
// steps. The drift of a step (position and orientation update, impl:135-147) is computed at
// the end of the previous step's velocity update; k_drift does it when no previous step
// left it. A rebuild is:
//   k_count           bucket key of every sphere and its count (posted atomic add)
//   k_scan_reduce     \
//   (k_scan_parts)     > exclusive prefix of the counts: start[b] = first sorted slot of
//   k_scan_apply      /  bucket b, start[nbucket] = N (tables of at most kDirectTiles tiles:
//                        each k_scan_apply block sums the earlier tile sums itself)
//   k_scatter         provisional slot start[key] + (atomic countdown of the bucket), which
//                     leaves every count at zero for the next step
//   k_order_gather    final slot = start[key] + number of spheres of the same bucket with a
//                     smaller id, i.e. exactly the reference's (hash, index) order; writes
//                     the sorted id and the FP32 position copy the pair search reads, with
//                     the sphere's wall-reach flag (FP64 test, reach + skin) in the sign of
//                     the radius; the wall gap of the sphere (k_touch); restarts the
//                     sphere's path length of the half-skin test
// The pair search k_pairs and the per-step pair test k_touch (xdem_step.hpp) are launched
// from here too.
// The bucket of slot range [start[b], start[b+1]) needs no clearing between steps, and a
// row of three consecutive buckets is the single range [start[b], start[b+3]).
// Built with -ffp-contract=off: the drift is bit-identical to the reference's.
#include <algorithm>

#include "xdem_step.hpp"

namespace xdem {
namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned kScanItems = 16;                   // per thread
constexpr unsigned kScanTile = kBlock * kScanItems;   // 4096 buckets per block
constexpr unsigned kPartsBlock = 1024;
constexpr unsigned kDirectTiles = 2048;

// Sum of x over the block (every thread gets it); red holds kBlock / 64 entries.
__device__ __forceinline__ uint32_t block_sum(uint32_t x, uint32_t* red)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (unsigned w = 0; w < kBlock / 64; w++) t += red[w];
    return t;
}

inline dim3 grid_for(unsigned n) { return dim3((n + kBlock - 1) / kBlock); }

// Drift of every sphere from the current state into the next position/orientation buffer
// (impl:135-147), with bucket keys and counts: used when the previous step did not leave its
// drift (the first step, or after the state was written from the host).
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_drift(StepArgs A)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.np) return;
    const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
    drift_sphere(A, S, i, ld4(A.pos + 4 * i), ld4(A.ep + 4 * i), ld3(A.vel + 3 * i), ld3(A.acc + 3 * i),
                 ld4(A.ev + 4 * i), ld4(A.ea + 4 * i));
}

// First kernel of every step: clears the step's contact counters, full-pass queue and the
// next step's rebuild flag. On a rebuild: the bucket key of every sphere and its count.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_count(StepArgs A)
{
    if (blockIdx.x == 0 && threadIdx.x < 2) A.counters[threadIdx.x] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x < kQueueSegs) {
        A.seg_count2[threadIdx.x * kSegStride] = 0u;
        A.dmax_next[threadIdx.x * kSegStride] = 0u;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *A.rebuild_next = 0u;
        ListCtl& L = *A.lctl;
        L.step += 1u;
        if (*A.rebuild_cur) {
            // skin of the lists this build makes, from how long the last ones lasted
            const uint32_t lasted = L.step - L.last_build;
            if (L.adaptive && L.last_build != 0u) {
                if (lasted < 16u) L.scale = fminf(1.0f, 2.0f * L.scale);
                else if (lasted > 64u) L.scale = fmaxf(L.min_scale, 0.5f * L.scale);
            }
            L.last_build = L.step;
        }
    }
    if (!*A.rebuild_cur) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) *A.n_rebuild += 1ull;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.np) return;
    const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
    const V4<double> p = ld4(A.pos + 4 * i);
    int gx, gy, gz;
    grid_cell(S.grid, p.x, p.y, p.z, gx, gy, gz);
    const uint32_t key = grid_bucket(S.grid, gx, gy, gz);
    A.keys[i] = key;
    atomicAdd(&A.bucket_count[key], 1u);   // result not waited for: k_scatter takes the rank
}

// Sum of 16 consecutive buckets per thread, reduced per block of 4096 buckets. Block 0 also
// clears the full-pass queue segment counters.
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ gate,
                                                        const uint32_t* __restrict__ count,
                                                        uint32_t* __restrict__ part,
                                                        uint32_t* __restrict__ seg_count)
{
    if (!*gate) return;
    if (blockIdx.x == 0 && threadIdx.x < kQueueSegs) seg_count[threadIdx.x * kSegStride] = 0u;
    __shared__ uint32_t wsum[kBlock / 64];
    const uint4* c = reinterpret_cast<const uint4*>(count + (size_t)blockIdx.x * kScanTile) + threadIdx.x * 4;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint4 v = c[k];
        s += v.x + v.y + v.z + v.w;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (unsigned w = 0; w < kBlock / 64; w++) t += wsum[w];
        part[blockIdx.x] = t;
    }
}

// Inclusive scan of one value per lane across a wave64.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

// Exclusive scan of the block sums by one block; writes the grand total to *total.
__global__ __launch_bounds__(kPartsBlock) void k_scan_parts(const uint32_t* __restrict__ gate,
                                                            uint32_t* __restrict__ part, uint32_t nparts,
                                                            uint32_t* __restrict__ total)
{
    if (!*gate) return;
    __shared__ uint32_t wsum[kPartsBlock / 64];
    const uint32_t per = (nparts + kPartsBlock - 1) / kPartsBlock;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t k = 0; k < per; k++)
        if (b0 + k < nparts) s += part[b0 + k];
    const uint32_t inc = wave_inclusive_scan(s);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (unsigned k = 0; k < kPartsBlock / 64; k++) {
            const uint32_t t = wsum[k];
            wsum[k] = run;
            run += t;
        }
        *total = run;
    }
    __syncthreads();
    uint32_t run = wsum[w] + inc - s;
    for (uint32_t k = 0; k < per; k++)
        if (b0 + k < nparts) {
            const uint32_t t = part[b0 + k];
            part[b0 + k] = run;
            run += t;
        }
}

// start[b] = (sum of the earlier tiles) + exclusive prefix inside the tile (the counts stay:
// k_scatter counts them back down to zero). DIRECT: part holds the tile sums and the block
// adds up those of the earlier tiles itself; the last block writes the total to *total.
// Otherwise part holds the exclusive prefix of the tile sums (k_scan_parts).
template <bool DIRECT>
__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t* __restrict__ gate,
                                                       const uint32_t* __restrict__ count,
                                                       const uint32_t* __restrict__ part,
                                                       uint32_t* __restrict__ start,
                                                       uint32_t* __restrict__ total)
{
    if (!*gate) return;
    __shared__ uint32_t wsum[kBlock / 64];
    __shared__ uint32_t red[kBlock / 64];
    uint32_t tile0;
    if (DIRECT) {
        uint32_t pre = 0;
        for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kBlock) pre += part[t];
        tile0 = block_sum(pre, red);
    } else {
        tile0 = part[blockIdx.x];
    }
    const size_t base = (size_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const uint4* c = reinterpret_cast<const uint4*>(count + base);
    uint32_t v[kScanItems];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint4 q = c[k];
        v[4 * k + 0] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < (int)kScanItems; k++) {
        const uint32_t t = v[k];
        v[k] = s;
        s += t;
    }
    const uint32_t inc = wave_inclusive_scan(s);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = tile0 + inc - s;
    for (uint32_t k = 0; k < w; k++) off += wsum[k];
    if (DIRECT && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) *total = off + s;
    uint4* o = reinterpret_cast<uint4*>(start + base);
#pragma unroll
    for (int k = 0; k < 4; k++)
        o[k] = make_uint4(off + v[4 * k], off + v[4 * k + 1], off + v[4 * k + 2], off + v[4 * k + 3]);
}

__global__ __launch_bounds__(kBlock) void k_scatter(StepArgs A)
{
    if (!*A.rebuild_cur) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.np) return;
    const uint32_t key = A.keys[i];
    // counting the bucket back down gives each sphere a distinct slot and leaves the count
    // at zero for the next step (the order inside the bucket is fixed by k_order_gather)
    A.slot_tmp[A.bucket_start[key] + atomicSub(&A.bucket_count[key], 1u) - 1u] = i;
}

template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_order_gather(StepArgs A)
{
    if (!*A.rebuild_cur) return;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.nsorted) return;
    const uint32_t k = A.slot_tmp[j];
    const uint32_t key = A.keys[k];
    const uint32_t s = A.bucket_start[key], e = A.bucket_start[key + 1];
    uint32_t r = 0;
    for (uint32_t m = s; m < e; m++) r += A.slot_tmp[m] < k ? 1u : 0u;
    const uint32_t f = s + r;
    A.svals[f] = k;
    const V4<double> p = ld4(A.pos + 4 * k);
    const SceneDev& S = A.scenes[scene_index(A, k, MULTI)];
    // exact FP64 reach test against every wall of the scene (wall_out_of_reach, radius
    // grown by the skin): a sphere with a wall in reach goes to the full contact pass; and
    // the least displacement that could bring a wall into reach, |dp . uw| / max(1, |uw|) -
    // r (1 + 1e-6), the gap k_touch compares with (any cohesive wall: none)
    bool far = true;
    double wgap = 3.0e38;
    for (uint32_t w = 0; w < S.nwall; w++) {
        const WallDev& W = A.walls[S.wall_begin + w];
        far = far & wall_out_of_reach(W.cc, W.info, p, S.skin * A.lctl->scale);
        const xdem_plane_info& pl = W.info;
        const double uw2 = pl.uw.x * pl.uw.x + pl.uw.y * pl.uw.y + pl.uw.z * pl.uw.z;
        const double wz = (p.x - pl.xw.x) * pl.uw.x + (p.y - pl.xw.y) * pl.uw.y + (p.z - pl.xw.z) * pl.uw.z;
        const double g = W.cc.coh != 0.0 ? -1.0 : fabs(wz) / sqrt(fmax(1.0, uw2)) - p.w * (1.0 + 1e-6) - 1e-9;
        wgap = fmin(wgap, g);
    }
    A.gap[k] = __double2float_rd(wgap);
    const Grid& gr = S.grid;
    const float rf = (float)p.w;
    const float4 q = make_float4((float)(p.x - gr.wo_x), (float)(p.y - gr.wo_y), (float)(p.z - gr.wo_z), rf);
    A.disp[k] = 0.0f;   // the path length of the half-skin test starts here
    A.sposf[f] = make_float4(q.x, q.y, q.z, far ? rf : -rf);
}

}  // namespace

uint32_t bucket_table_size(uint32_t nbucket)
{
    return (nbucket + kScanTile - 1) / kScanTile * kScanTile;
}

hipError_t launch_drift(const StepArgs& A, bool multi, hipStream_t s)
{
    if (A.np == 0) return hipSuccess;
    if (multi) hipLaunchKernelGGL(k_drift<true>, grid_for(A.np), dim3(kBlock), 0, s, A);
    else hipLaunchKernelGGL(k_drift<false>, grid_for(A.np), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_pairs(const StepArgs& A, bool multi, hipStream_t s)
{
    if (A.nsorted == 0) return hipSuccess;
    if (multi) hipLaunchKernelGGL(k_pairs<true>, grid_for(A.nsorted), dim3(kBlock), 0, s, A);
    else hipLaunchKernelGGL(k_pairs<false>, grid_for(A.nsorted), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_touch(const StepArgs& A, bool multi, hipStream_t s)
{
    const dim3 g(kQueueSegs * kQueueChunks);
    if (multi) hipLaunchKernelGGL(k_touch<true>, g, dim3(kBlock), 0, s, A);
    else hipLaunchKernelGGL(k_touch<false>, g, dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_grid(const StepArgs& A, bool multi, hipStream_t s)
{
    const uint32_t nb = bucket_table_size(A.nbucket);
    const uint32_t ntile = nb / kScanTile;
    const uint32_t* gate = A.rebuild_cur;
    // the first kernel also runs when np is 0: it clears the step's counters and flag
    if (multi) hipLaunchKernelGGL(k_count<true>, grid_for(std::max(A.np, 1u)), dim3(kBlock), 0, s, A);
    else hipLaunchKernelGGL(k_count<false>, grid_for(std::max(A.np, 1u)), dim3(kBlock), 0, s, A);
    hipLaunchKernelGGL(k_scan_reduce, dim3(ntile), dim3(kBlock), 0, s, gate, A.bucket_count, A.scan_part,
                       A.seg_count);
    if (ntile <= kDirectTiles) {
        hipLaunchKernelGGL(k_scan_apply<true>, dim3(ntile), dim3(kBlock), 0, s, gate, A.bucket_count, A.scan_part,
                           A.bucket_start, A.bucket_start + nb);
    } else {
        hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(kPartsBlock), 0, s, gate, A.scan_part, ntile,
                           A.bucket_start + nb);
        hipLaunchKernelGGL(k_scan_apply<false>, dim3(ntile), dim3(kBlock), 0, s, gate, A.bucket_count, A.scan_part,
                           A.bucket_start, A.bucket_start + nb);
    }
    if (A.np) {
        hipLaunchKernelGGL(k_scatter, grid_for(A.np), dim3(kBlock), 0, s, A);
        if (multi) hipLaunchKernelGGL(k_order_gather<true>, grid_for(A.nsorted), dim3(kBlock), 0, s, A);
        else hipLaunchKernelGGL(k_order_gather<false>, grid_for(A.nsorted), dim3(kBlock), 0, s, A);
    }
    return hipGetLastError();
}

}  // namespace xdem

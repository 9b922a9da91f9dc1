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
// left it. A rebuild is one kernel, k_rebuild, in phases (each waits until the previous one
// is complete; the kernel's first lines, on every step, clear the step's counters, and the
// steps that do not rebuild return right after them):
//   count             bucket key of every sphere and its count (posted atomic add)
//   scan              exclusive prefix of the counts, tile sums then tiles: start[b] =
//                     first sorted slot of bucket b, start[nbucket] = N (each tile adds up
//                     the earlier tile sums itself up to kDirectTiles tiles, else one block
//                     scans them first)
//   scatter           provisional slot start[key] + (atomic countdown of the bucket), which
//                     leaves every count at zero for the next build
//   order + gather    final slot = start[key] + number of spheres of the same bucket with a
//                     smaller id, i.e. exactly the reference's (hash, index) order; writes
//                     the sorted id and the FP32 position copy the pair search reads, with
//                     the sphere's wall-reach flag (FP64 test, reach + skin) in the sign of
//                     the radius; the wall gap of the sphere (k_touch); restarts the
//                     sphere's path length of the half-skin test
// The per-step pair test k_touch (xdem_step.hpp), which also does the pair search on a
// build step, is launched from here too.
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

// k_rebuild's phases hand data to blocks on other XCDs, whose L2 caches are not coherent
// with each other. Everything one phase writes for a later one is stored and loaded with
// agent-scope atomic accesses (written through to memory, read from it), so the wait
// between phases needs no cache write-back or invalidation. Loads of data a phase may have left stale are
// also checked against the array bounds (XDEM_ERR_GRID_SYNC instead of a wild access).
__device__ __forceinline__ uint32_t ld_co(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_co(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every step, block 0: clears the step's contact counters, full-pass queue and the next
// step's rebuild flag.
__device__ __forceinline__ void step_counters(const StepArgs& A)
{
    if (blockIdx.x != 0) return;
    if (threadIdx.x < 2) A.counters[threadIdx.x] = 0ull;
    if (threadIdx.x < kQueueSegs) {
        A.seg_count2[threadIdx.x * kSegStride] = 0u;
        A.dmax_next[threadIdx.x * kSegStride] = 0u;
    }
    if (threadIdx.x == 0) *A.rebuild_next = 0u;
}

// Every step, one thread: counts the step for the adaptive skin; on a rebuild, sets the skin
// of the lists it makes from how long the last ones lasted (read by the order phase).
__device__ __forceinline__ void list_ctl_update(const StepArgs& A, bool rebuild)
{
    ListCtl& L = *A.lctl;
    L.step += 1u;
    if (!rebuild) return;
    const uint32_t lasted = L.step - L.last_build;
    if (L.adaptive && L.last_build != 0u) {
        float sc = L.scale;
        if (lasted < 16u) sc = fminf(1.0f, 2.0f * sc);
        else if (lasted > 64u) sc = fmaxf(L.min_scale, 0.5f * sc);
        st_co(reinterpret_cast<uint32_t*>(&L.scale), __float_as_uint(sc));
    }
    L.last_build = L.step;
    L.body_path_build = L.body_path;   // the lists of this build see the bodies where they are
    *A.n_rebuild += 1ull;
}

// The phases of a rebuild hand out their work in units (kUnit spheres, or one scan tile)
// from atomic cursors (kClaimGroups ranges of units per phase, each with its own cursor);
// a block claims units until none is left, then waits for the phase to be complete: every
// unit done, counted per range, and the completed ranges counted per phase. The block that
// finishes the last unit writes the phase's number to kBarFlags release flags
// (one per 128-byte line, one wave store), and each block waits on flag blockIdx mod
// kBarFlags. A block only ever waits for units that running blocks have claimed, so the
// phases complete whether or not every block of the grid is resident (other kernels or
// other contexts may hold the device). Every wave waits for its stores to complete (they are
// written through) before its unit is counted. k_touch, the next kernel of the step, sets the
// cursors, counters and flags back to zero. The wait is bounded (kSyncSpins tries of ~1 us,
// or the context's rebuild_wait_limit): beyond it the block sets XDEM_ERR_GRID_SYNC (a
// warning) and leaves the kernel, so it always ends. The build still completes: the block
// that finishes a phase's last unit has released the phase before it waits, so it passes and
// claims the next phase's units, and every unit of every phase is claimed by a block that
// stays. Every index a later phase reads from an earlier one is bounds-checked all the same.
constexpr uint32_t kSyncSpins = 1u << 20;
constexpr uint32_t kUnit = 4 * kBlock;

template <class F>
__device__ __forceinline__ bool run_phase(const StepArgs& A, uint32_t k, uint32_t nunits, F&& body)
{
    if (nunits == 0) return true;
    __shared__ uint32_t claim, passed;
    // kClaimGroups equal ranges of units, each with its own cursor and done counter (fewer
    // atomics per address); a block starts on range blockIdx mod kClaimGroups and moves on to
    // the next range once one is exhausted
    uint32_t nonempty = 0;
#pragma unroll
    for (uint32_t c = 0; c < kClaimGroups; c++) nonempty += (c + 1u) * nunits / kClaimGroups > c * nunits / kClaimGroups;
    uint32_t c = blockIdx.x % kClaimGroups, tried = 0;
    while (tried < kClaimGroups) {
        const uint32_t lo = c * nunits / kClaimGroups, hi = (c + 1u) * nunits / kClaimGroups;
        __syncthreads();   // every thread has read the last claim
        if (threadIdx.x == 0) claim = lo < hi ? lo + atomicAdd(A.grid_bar + kSegStride * bar_cursor(k, c), 1u) : hi;
        __syncthreads();
        const uint32_t u = claim;
        if (u >= hi) {
            c = (c + 1u) % kClaimGroups;
            tried++;
            continue;
        }
        body(u);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (threadIdx.x < 64) {
            uint32_t last = 0;
            if (threadIdx.x == 0 && atomicAdd(A.grid_bar + kSegStride * bar_done(k, c), 1u) + 1u == hi - lo)
                last = atomicAdd(A.grid_bar + kSegStride * bar_groups(k), 1u) + 1u == nonempty ? 1u : 0u;
            last = __shfl(last, 0, 64);
            if (last) st_co(A.grid_bar + kSegStride * bar_flag(threadIdx.x), k + 1u);
        }
    }
    if (threadIdx.x == 0) {
        const uint32_t* flag = A.grid_bar + kSegStride * bar_flag(blockIdx.x % kBarFlags);
        uint32_t tries = 0;
        bool ok = false;
        const uint32_t limit = A.rebuild_wait_limit ? A.rebuild_wait_limit : kSyncSpins;
        while (true) {
            ok = ld_co(flag) >= k + 1u;
            if (ok || ++tries >= limit) break;
            __builtin_amdgcn_s_sleep(2);
        }
        if (!ok) atomicOr(A.err, XDEM_ERR_GRID_SYNC);
        passed = ok ? 1u : 0u;
    }
    __syncthreads();
    return passed != 0u;
}

// Sum of tile t (4096 buckets) into part[t]; lane-contiguous loads (coalesced).
__device__ __forceinline__ void scan_reduce_tile(const uint32_t* count, uint32_t* part, uint32_t t)
{
    __shared__ uint32_t wsum[kBlock / 64];
    const uint32_t* c = count + (size_t)t * kScanTile + threadIdx.x;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < (int)kScanItems; k++) s += ld_co(c + k * kBlock);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t u = 0;
#pragma unroll
        for (unsigned w = 0; w < kBlock / 64; w++) u += wsum[w];
        st_co(part + t, u);
    }
    __syncthreads();
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

// Exclusive scan of the tile sums by one block (tables of more than kDirectTiles tiles);
// writes the grand total to *total.
__device__ __forceinline__ void scan_parts_block(uint32_t* part, uint32_t nparts, uint32_t* total)
{
    __shared__ uint32_t wsum[kBlock / 64];
    const uint32_t per = (nparts + kBlock - 1) / kBlock;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t k = 0; k < per; k++)
        if (b0 + k < nparts) s += ld_co(part + b0 + k);
    const uint32_t inc = wave_inclusive_scan(s);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (unsigned k = 0; k < kBlock / 64; k++) {
            const uint32_t t = wsum[k];
            wsum[k] = run;
            run += t;
        }
        st_co(total, run);
    }
    __syncthreads();
    uint32_t run = wsum[w] + inc - s;
    for (uint32_t k = 0; k < per; k++)
        if (b0 + k < nparts) {
            const uint32_t t = ld_co(part + b0 + k);
            st_co(part + b0 + k, run);
            run += t;
        }
    __syncthreads();
}

// start[b] of tile t = (sum of the earlier tiles) + exclusive prefix inside the tile (the
// counts stay: the scatter counts them back down to zero). DIRECT: part holds the tile sums
// and the block adds up those of the earlier tiles itself; the last tile writes the total
// to *total. Otherwise part holds the exclusive prefix of the tile sums (scan_parts_block).
template <bool DIRECT>
__device__ __forceinline__ void scan_apply_tile(const uint32_t* count, const uint32_t* part, uint32_t* start,
                                                uint32_t* total, uint32_t t, uint32_t ntile, uint32_t* tile)
{
    __shared__ uint32_t wsum[kBlock / 64];
    __shared__ uint32_t red[kBlock / 64];
    uint32_t tile0;
    if (DIRECT) {
        uint32_t pre = 0;
        for (uint32_t u = threadIdx.x; u < t; u += kBlock) pre += ld_co(part + u);
        tile0 = block_sum(pre, red);
    } else {
        tile0 = ld_co(part + t);
    }
    // the tile through LDS (tile: kScanTile + kScanTile / kScanItems words): coalesced loads
    // and stores, kScanItems consecutive buckets per thread for the serial part (one pad
    // word per kScanItems: conflict-free rows)
    const size_t base = (size_t)t * kScanTile;
#pragma unroll
    for (int k = 0; k < (int)kScanItems; k++) {
        const uint32_t q = k * kBlock + threadIdx.x;
        tile[q + q / kScanItems] = ld_co(count + base + q);
    }
    __syncthreads();
    uint32_t* row = tile + threadIdx.x * (kScanItems + 1);
    uint32_t v[kScanItems];
#pragma unroll
    for (int k = 0; k < (int)kScanItems; k++) v[k] = row[k];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < (int)kScanItems; k++) {
        const uint32_t u = v[k];
        v[k] = s;
        s += u;
    }
    const uint32_t inc = wave_inclusive_scan(s);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = tile0 + inc - s;
    for (uint32_t k = 0; k < w; k++) off += wsum[k];
    if (DIRECT && t == ntile - 1 && threadIdx.x == kBlock - 1) st_co(total, off + s);
#pragma unroll
    for (int k = 0; k < (int)kScanItems; k++) row[k] = off + v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < (int)kScanItems; k++) {
        const uint32_t q = k * kBlock + threadIdx.x;
        st_co(start + base + q, tile[q + q / kScanItems]);
    }
    __syncthreads();
}

// Count phase unit u: spheres u * kUnit + r * kBlock + thread, r < 4, their positions read
// together.
template <bool MULTI>
__device__ __forceinline__ void count_unit(const StepArgs& A, uint32_t u)
{
    V4<double> p[4];
    uint32_t idx[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        idx[r] = u * kUnit + r * kBlock + threadIdx.x;
        p[r] = ld4(A.pos + 4 * (size_t)min(idx[r], A.np - 1u));
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        if (idx[r] >= A.np) continue;
        const SceneDev& S = A.scenes[scene_index(A, idx[r], MULTI)];
        int gx, gy, gz;
        grid_cell(S.grid, p[r].x, p[r].y, p[r].z, gx, gy, gz);
        const uint32_t key = grid_bucket(S.grid, gx, gy, gz);
        st_co(A.keys + idx[r], key);
        atomicAdd(&A.bucket_count[key], 1u);   // result not waited for: the scatter takes the rank
    }
}

// Provisional slot of sphere i: counting the bucket back down gives each sphere a distinct
// slot and leaves the count at zero for the next build (the order inside the bucket is
// fixed by order_gather_slot).
__device__ __forceinline__ void scatter_sphere(const StepArgs& A, uint32_t i)
{
    const uint32_t key = ld_co(A.keys + i);
    if (key >= A.nbucket) { atomicOr(A.err, XDEM_ERR_GRID_SYNC); return; }
    const uint32_t slot = ld_co(A.bucket_start + key) + atomicSub(&A.bucket_count[key], 1u) - 1u;
    if (slot < A.np) st_co(A.slot_tmp + slot, i);
    else atomicOr(A.err, XDEM_ERR_GRID_SYNC);
}

template <bool MULTI, bool CYL>
__device__ __forceinline__ void order_gather_slot(const StepArgs& A, uint32_t j)
{
    const uint32_t k = ld_co(A.slot_tmp + j);
    if (k >= A.np) { atomicOr(A.err, XDEM_ERR_GRID_SYNC); return; }
    const uint32_t key = ld_co(A.keys + k);
    if (key >= A.nbucket) { atomicOr(A.err, XDEM_ERR_GRID_SYNC); return; }
    const uint32_t s = ld_co(A.bucket_start + key), e = min(ld_co(A.bucket_start + key + 1), A.nsorted);
    uint32_t r = 0;
    for (uint32_t m = s; m < e; m++) r += ld_co(A.slot_tmp + m) < k ? 1u : 0u;
    const uint32_t f = s + r;
    if (f >= A.nsorted) { atomicOr(A.err, XDEM_ERR_GRID_SYNC); return; }
    A.svals[f] = k;
    const V4<double> p = ld4(A.pos + 4 * k);
    const SceneDev& S = A.scenes[scene_index(A, k, MULTI)];
    // exact FP64 reach test against every wall and cylinder of the scene (radius grown by the
    // skin): a sphere with one in reach goes to the full contact pass; and
    // the least displacement that could bring a wall into reach, |dp . uw| / max(1, |uw|) -
    // r (1 + 1e-6), the gap k_touch compares with (any cohesive wall: none)
    const float scale = __uint_as_float(ld_co(reinterpret_cast<const uint32_t*>(&A.lctl->scale)));
    bool far = true;
    double wgap = 3.0e38;
    for (uint32_t w = 0; w < S.nwall; w++) {
        const WallDev& W = A.walls[S.wall_begin + w];
        far = far & wall_out_of_reach(W.cc, W.info, p, S.skin * scale);
        const xdem_plane_info& pl = W.info;
        const double uw2 = pl.uw.x * pl.uw.x + pl.uw.y * pl.uw.y + pl.uw.z * pl.uw.z;
        const double wz = (p.x - pl.xw.x) * pl.uw.x + (p.y - pl.xw.y) * pl.uw.y + (p.z - pl.xw.z) * pl.uw.z;
        const double g = W.cc.coh != 0.0 ? -1.0 : fabs(wz) / sqrt(fmax(1.0, uw2)) - p.w * (1.0 + 1e-6) - 1e-9;
        wgap = fmin(wgap, g);
    }
    // cylinders: in reach with the radius grown by the skin, and the gap the largest overlap
    // lacks (it changes by at most the distance the sphere or the body moves)
    for (uint32_t c = 0; CYL && c < S.ncyl; c++) {
        const CylDev& C = A.cyls[S.cyl_begin + c];
        far = far & cylinder_out_of_reach(C.cc, C.info, C.dbi, C.margin, p, S.skin * scale);
        const double g = C.cc.coh != 0.0 ? -1.0
                                         : -cylinder_reach(C.info, cylinder_axis(C.info, C.dbi), mk3(p.x, p.y, p.z),
                                                           p.w * (1.0 + 1e-6)) - C.margin - 1e-9;
        wgap = fmin(wgap, g);
    }
    A.gap[k] = __double2float_rd(wgap);
    const Grid& gr = S.grid;
    const float rf = (float)p.w;
    const float4 q = make_float4((float)(p.x - gr.wo_x), (float)(p.y - gr.wo_y), (float)(p.z - gr.wo_z), rf);
    A.disp[k] = 0.0f;   // the path length of the half-skin test starts here
    A.sposf[f] = make_float4(q.x, q.y, q.z, far ? rf : -rf);
}

// First kernel of every step: the step's counters (block 0), then, on a rebuild only, the
// bucket grid in phases (work units claimed from per-phase cursors; the other steps return
// at once, so a step that does not rebuild costs one launch, not six).
template <bool MULTI, bool CYL>
__global__ __launch_bounds__(kBlock) void k_rebuild(StepArgs A, uint32_t ntile)
{
    step_counters(A);
    const bool rebuild = *A.rebuild_cur != 0u;
    if (!rebuild) {
        if (blockIdx.x == 0 && threadIdx.x == 0) list_ctl_update(A, false);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x < kQueueSegs) A.seg_count[threadIdx.x * kSegStride] = 0u;   // candidate queue
    if (blockIdx.x == 0 && threadIdx.x == 0 && A.has_mesh) *A.mesh_count = 0u;                      // mesh queue
    const uint32_t sp_units = (A.np + kUnit - 1) / kUnit;
    if (!run_phase(A, 0, sp_units, [&](uint32_t u) {
            if (u == 0 && threadIdx.x == 0) list_ctl_update(A, true);   // before the order phase reads the skin
            count_unit<MULTI>(A, u);
        }))
        return;
    if (!run_phase(A, 1, ntile, [&](uint32_t t) { scan_reduce_tile(A.bucket_count, A.scan_part, t); })) return;
    __shared__ uint32_t tile[kScanTile + kScanTile / kScanItems];
    uint32_t* total = A.bucket_start + (size_t)ntile * kScanTile;
    const bool direct = ntile <= kDirectTiles;
    if (!run_phase(A, 2, direct ? 0u : 1u, [&](uint32_t) { scan_parts_block(A.scan_part, ntile, total); })) return;
    if (!run_phase(A, 3, ntile, [&](uint32_t t) {
            if (direct) scan_apply_tile<true>(A.bucket_count, A.scan_part, A.bucket_start, total, t, ntile, tile);
            else scan_apply_tile<false>(A.bucket_count, A.scan_part, A.bucket_start, total, t, ntile, tile);
        }))
        return;
    if (!run_phase(A, 4, sp_units, [&](uint32_t u) {
            for (uint32_t r = 0; r < 4; r++) {
                const uint32_t i = u * kUnit + r * kBlock + threadIdx.x;
                if (i < A.np) scatter_sphere(A, i);
            }
        }))
        return;
    const uint32_t sl_units = (A.nsorted + kUnit - 1) / kUnit;
    run_phase(A, 5, sl_units, [&](uint32_t u) {
        for (uint32_t r = 0; r < 4; r++) {
            const uint32_t j = u * kUnit + r * kBlock + threadIdx.x;
            if (j < A.nsorted) order_gather_slot<MULTI, CYL>(A, j);
        }
    });
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

hipError_t launch_touch(const StepArgs& A, bool multi, hipStream_t s)
{
    // the resident grid: a build step loops over the sorted slots, the other steps stride
    // over the candidate queue's segments
    const dim3 g(A.resident_blocks);
    if (multi) {
        if (A.has_cyl) hipLaunchKernelGGL((k_touch<true, true>), g, dim3(kBlock), 0, s, A);
        else hipLaunchKernelGGL((k_touch<true, false>), g, dim3(kBlock), 0, s, A);
    } else {
        if (A.has_cyl) hipLaunchKernelGGL((k_touch<false, true>), g, dim3(kBlock), 0, s, A);
        else hipLaunchKernelGGL((k_touch<false, false>), g, dim3(kBlock), 0, s, A);
    }
    return hipGetLastError();
}

hipError_t launch_grid(const StepArgs& A, bool multi, hipStream_t s)
{
    const uint32_t ntile = bucket_table_size(A.nbucket) / kScanTile;
    if (A.rebuild_blocks == 0) return hipErrorInvalidValue;
    const dim3 g(A.rebuild_blocks);
    if (multi) {
        if (A.has_cyl) hipLaunchKernelGGL((k_rebuild<true, true>), g, dim3(kBlock), 0, s, A, ntile);
        else hipLaunchKernelGGL((k_rebuild<true, false>), g, dim3(kBlock), 0, s, A, ntile);
    } else {
        if (A.has_cyl) hipLaunchKernelGGL((k_rebuild<false, true>), g, dim3(kBlock), 0, s, A, ntile);
        else hipLaunchKernelGGL((k_rebuild<false, false>), g, dim3(kBlock), 0, s, A, ntile);
    }
    return hipGetLastError();
}

uint32_t rebuild_grid_size(bool multi)
{
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    // the smaller residency of the two instantiations (with and without cylinders)
    int per2 = 0;
    hipError_t e = multi ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rebuild<true, false>, kBlock, 0)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rebuild<false, false>, kBlock, 0);
    if (e == hipSuccess)
        e = multi ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per2, k_rebuild<true, true>, kBlock, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per2, k_rebuild<false, true>, kBlock, 0);
    per = std::min(per, per2);
    if (e != hipSuccess || ncu <= 0 || per <= 0) return 0;
    // two blocks per CU (the phases do not need them all resident: work goes to the blocks
    // that run; more blocks only add claims on the phase cursors: measured slower)
    const uint32_t g = std::min((uint32_t)ncu * (uint32_t)per, 2u * (uint32_t)ncu);
    return g;
}

}  // namespace xdem

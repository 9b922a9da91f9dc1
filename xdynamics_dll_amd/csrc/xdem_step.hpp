// xdem_step.hpp -- kernels of the fused DEM step (the xdem_ctx path).
//
// One step is the grid build of xdem_grid.hip (position update, bucket counting sort,
// sorted copies) followed by
//   k_collide         one thread per sorted slot: sphere-sphere contacts over the 27
//                     neighbour buckets in the reference's (z, y, x) order, every wall of
//                     the scene in id order, then (fused mode) rolling resistance and the
//                     velocity/orientation-rate update (xIntegratorVV.cpp:131-152);
//   k_finish          (split mode) rolling resistance and the velocity update.
// Several independent scenes share the launches: each scene owns a range of buckets, so
// spheres of different scenes never meet in a bucket.
#pragma once

#include "xdem_contact.hpp"

namespace xdem {

constexpr uint32_t kInvalid = 0xffffffffu;
// nbr_count of a sphere: the length of its neighbour list and three flags
constexpr uint32_t kCandidate = 0x80000000u;  // listed neighbours or a wall within reach + skin (build)
constexpr uint32_t kWallNear = 0x40000000u;   // a wall within reach + skin (build)
constexpr uint32_t kTouch = 0x20000000u;      // this step: a pair passes or a wall is in reach (k_touch)
constexpr uint32_t kSpinStable = 0x10000000u; // both ev and ep buffers hold the sphere's current values
constexpr uint32_t kAccStable = 0x08000000u;  // acc holds the sphere's contact-free acceleration
constexpr uint32_t kNoContacts = 0x04000000u; // both p2p count buffers, the wall count and the mesh
                                              // count hold 0
constexpr uint32_t kMeshNear = 0x02000000u;   // triangles of the scene's mesh within reach + skin (build);
                                              // such a sphere goes to the full pass every step
constexpr uint32_t kLiteStable = kSpinStable | kAccStable | kNoContacts;   // set by the lite pass
constexpr uint32_t kCountMask = 0x01ffffffu;
// The candidate and full-pass queues are split in kQueueSegs segments, each with its own
// counter on its own 128-byte line: the wave of sorted slots [64 w, 64 w + 64) appends to
// segment w mod kQueueSegs, so the appends of different waves rarely hit the same counter and
// a contact region spreads over every segment. (Segments of 64 contiguous ranges of slots,
// each XCD walking adjacent ones, made a settled bed's step 6% faster and the dam break's 2-3%
// slower: its few thousand full-pass spheres then fell into a few segments, on few XCDs.)
constexpr uint32_t kQueueSegs = 64;
constexpr uint32_t kSegStride = 32;           // uint32 entries between two segment counters
constexpr uint32_t kBarFlags = 64;            // release flags of k_rebuild's phases
constexpr uint32_t kRebuildPhases = 6;        // k_rebuild: count, reduce, parts, apply, scatter, order
constexpr uint32_t kClaimGroups = 8;          // unit ranges per phase, each with a cursor and a counter
// k_rebuild's slots (each on its own 128-byte line): per phase and range a cursor and a done
// counter, per phase a count of completed ranges, and the release flags (each holds the
// number of the last completed phase + 1; kRebuildPhases once the build is complete)
__host__ __device__ constexpr uint32_t bar_cursor(uint32_t k, uint32_t c) { return k * kClaimGroups + c; }
__host__ __device__ constexpr uint32_t bar_done(uint32_t k, uint32_t c) { return (kRebuildPhases + k) * kClaimGroups + c; }
__host__ __device__ constexpr uint32_t bar_groups(uint32_t k) { return 2 * kRebuildPhases * kClaimGroups + k; }
__host__ __device__ constexpr uint32_t bar_flag(uint32_t f) { return 2 * kRebuildPhases * kClaimGroups + kRebuildPhases + f; }
constexpr uint32_t kBarSlots = bar_flag(kBarFlags);
// Entries per segment for nsorted sorted slots launched in blocks of 256 (the most a
// segment can receive: 64 spheres from each of its waves).
inline uint32_t queue_segment_cap(uint32_t nsorted)
{
    const uint32_t nwaves = (nsorted + 255u) / 256u * 4u;
    return (nwaves + kQueueSegs - 1u) / kQueueSegs * 64u;
}
constexpr int kMaxWallsPerScene = 16;
constexpr uint32_t kTriCap = 16;              // triangles listed per sphere (mesh contacts)
constexpr uint32_t kWrenchBlocks = 256;       // partial sums per body of the wrench reduction

// Neighbour-list skin control (device memory, one per context). Every step counts `step`;
// on a list build with `adaptive` set, the skin scale doubles when the lists lasted fewer
// than 16 steps and halves when they lasted more than 64 (within [min_scale, 1]). The scale
// changes only at the start of a build, before anything uses it, so each list generation
// has one skin throughout. body_path is the summed displacement bound of every pose update of
// a wall or mesh body (each update adds the largest distance any point of the body moved),
// body_path_build its value at the last build: the bodies have moved at most their difference
// since, and a pose update that makes it reach half the smallest skin asks for a build.
struct ListCtl {
    float scale, min_scale;
    uint32_t step, last_build, adaptive, pad_;
    double body_path, body_path_build;
    double half_skin_min;     // half the smallest scene skin at scale 1
};

struct SceneDev {
    double dt, half2dt;
    double gx, gy, gz;
    Grid grid;
    ContactConst cc;          // sphere-sphere material pair
    uint32_t p_begin, p_end;  // sphere range of the scene
    uint32_t wall_begin, nwall;
    uint32_t has_pp;
    uint32_t wall_id0;        // a wall of the scene has plane id 0 (it resets the wall slot count)
    uint32_t wall_dynamic;    // a wall of the scene is a dynamic body (per-sphere wrench output)
    uint32_t pad_;
    double skin;              // largest neighbour-list skin (0: the lists are rebuilt every step);
                              // the lists use skin * ListCtl::scale
    double ref_cs;            // the reference's cell size (visiting order of the contacts);
                              // grid.cs, the bucket size, is larger by the skin and reach
    int32_t mesh;             // the scene's mesh body (index into StepArgs::meshes), -1 for none
    uint32_t cyl_begin;       // the scene's cylinder bodies in StepArgs::cyls
    uint32_t ncyl;
    uint32_t cyl_id0;         // a cylinder of the scene has id 0 (it restarts the cylinder slots)
};

// A rigid triangle mesh body of a scene. Triangles keep their body-frame vertices; the
// global ones (and the proxy sphere centres of the reference's cell binning) are recomputed
// from them at every pose. A static grid in the body frame (cells of side h over the mesh's
// local bounding box) lists every triangle within reach of any point of a cell: a sphere
// centre's cell holds every triangle its list can need.
struct MeshDev {
    double dbi[16];           // pose MM..ED3 (15 used)
    ContactConst cc;          // sphere-mesh material pair
    uint32_t tri_begin, ntri; // triangles in the context's triangle arrays
    int32_t wslot;            // per-sphere wrench slot of a dynamic mesh, -1 for a fixed one
    uint32_t cell_base;       // first cell of the body-frame grid in the context's cell table
    double lo_x, lo_y, lo_z, h;
    uint32_t dim_x, dim_y, dim_z, pad_;
};

struct WallDev {
    xdem_plane_info info;
    double dbi[16];           // 15 used
    ContactConst cc;
    int32_t wslot;            // per-sphere wrench slot of a dynamic wall, -1 for a fixed wall
    uint32_t pad_;
};

// A cylinder body of a scene (the reference's device_cylinder_info and its body pose).
struct CylDev {
    xdem_cylinder_info info;
    double dbi[16];           // 15 used
    ContactConst cc;
    double margin;            // widens the reach test where the contact regions do not join
                              // continuously (side and rim radii that differ, a lid ring)
    int32_t wslot;            // per-sphere wrench slot of a dynamic cylinder, -1 for a fixed one
    uint32_t pad_;
};

struct StepArgs {
    // state (reference layouts, all scenes concatenated)
    double* pos;  double* vel;  double* acc;  double* ep;  double* ev;  double* ea;
    double* force; double* moment; const double* mass; const double* inertia;
    double* tmax; double* rres;
    double* vel_out; double* ev_out;  // where the fused contact kernel writes vel and ev (the
                                      // other buffer of a pair: neighbours read the old ones)
    double* pos_next; double* ep_next; // the next step's drifted position and orientation
    int drift;                        // the velocity update also computes the next drift
    const uint32_t* scene_of;  // sphere -> scene (multi-scene only)
    const SceneDev* scenes;
    const WallDev* walls;
    const CylDev* cyls;
    uint32_t* cy_count; uint32_t* cy_id; double2* cy_tsd;   // cylinder slots (3 per sphere, in place)
    // grid: bucket key per sphere; dense bucket table
    uint32_t* keys;
    uint32_t* bucket_count;                  // zero between steps
    uint32_t* bucket_start;                  // first sorted slot of each bucket, [nbucket] = N
    uint32_t* scan_part;                     // per 4096-bucket tile sums / offsets
    uint32_t* slot_tmp;                      // sphere id per provisional slot
    uint32_t* svals;                         // sphere id per sorted slot
    uint32_t* work;                          // candidate spheres (pair search), by queue segment
    uint32_t* work2;                         // spheres of this step's full contact pass (k_touch)
    unsigned long long* pass_mask;           // per sphere of work2: list entries that pass (k_touch)
    const uint32_t* ext_id;                  // per stored sphere: the caller's (global) index
    const uint32_t* int_id;                  // per caller's (global) index: the stored sphere
    const double* jflat;                     // per stored sphere: quirk Q5's rolling inertia
    uint32_t* nbr_count;                     // listed neighbours per sphere | kCandidate, kWallNear, kTouch,
                                             // kSpinStable, kAccStable, kNoContacts
    uint32_t* nbr;                           // their sphere ids, list_cap per sphere
    float4* sposf;  // per sorted slot: position relative to the scene's grid origin and
                    // radius, FP32 (the pair search reads only this); the radius is
                    // negated when a wall is within reach (+ skin) of the sphere
    uint32_t* rebuild_cur;                   // nonzero: this step rebuilds grid and lists
    uint32_t* rebuild_next;                  // set when a sphere moved half the skin (or null)
    ListCtl* lctl;                           // skin scale of the current lists
    float* gap;                              // per sphere: lower bound of the distance its listed
                                             // pairs and near walls lack to pass (last build)
    float* disp;                             // per sphere: upper bound of its displacement since
                                             // the last build (path length, summed by the drift)
    uint32_t* dmax_cur;                      // largest disp of this step (bits of a float, per
    uint32_t* dmax_next;                     // queue segment); the drift writes the next one's
    unsigned long long* n_rebuild;           // list builds so far
    // sphere history (double-buffered) and wall history (in place)
    const uint32_t* pp_count_in; const uint32_t* pp_id_in; const double2* pp_tsd_in;
    uint32_t* pp_count_out; uint32_t* pp_id_out; double2* pp_tsd_out;
    uint32_t* pl_count; uint32_t* pl_id; double2* pl_tsd;
    double* wrench;          // [nwslot][6][np] per-sphere contributions to dynamic walls and meshes
    // meshes
    const MeshDev* meshes;
    const xdem_triangle_info* tri_glob;      // global triangles at the current pose
    const double* prox_glob;                 // [4 T] global proxy sphere centres (+ radius)
    const uint32_t* mcell_start;             // body-frame grid: triangles of cell c are
    const uint32_t* mcell_tris;              // mcell_tris[mcell_start[c] .. mcell_start[c + 1])
    uint32_t* tri_count;                     // per sphere: listed triangles (build)
    uint32_t* tri_list;                      // their indices in the mesh, kTriCap per sphere
    double* msum;                            // per sphere: triangle sums F, M, T (3 each), R
    double* ppsum;                           // per sphere of the full pass: sphere-contact sums F, M, T, R
    uint32_t* ms_count; uint32_t* ms_id; double2* ms_tsd;   // triangle history, 5 slots (in place)
    uint32_t* mesh_work;                     // spheres with listed triangles (build), mesh_count of them
    uint32_t* mesh_count;
    int has_mesh;
    uint32_t* err;
    unsigned long long* counters;  // [0] stored sphere contacts, [1] stored wall contacts
    uint32_t* seg_count;           // spheres per candidate queue segment (kSegStride apart)
    uint32_t* seg_count2;          // spheres per full-pass queue segment (kSegStride apart)
    uint32_t seg_cap;              // entries per queue segment in work and work2
    uint32_t* grid_bar;            // k_rebuild's phase cursors, done counters and release flags,
                                   // kBarSlots, each on its own 128-byte line (zero between steps)
    uint32_t rebuild_blocks;       // k_rebuild's grid
    uint32_t rebuild_wait_limit;   // tries before a k_rebuild block gives up waiting on a phase
    uint32_t np, nsorted, cap;
    uint32_t nwslot, nbucket;
    uint32_t list_cap;                       // neighbour-list entries per sorted slot
    uint32_t full_blocks;                    // grid of the full-pass kernels (host side: launch shape)
    uint32_t resident_blocks;                // grid of k_touch and k_pairs: the blocks the device holds
                                             // at once at three waves per SIMD (host side)
    uint32_t merge_below;                    // full passes smaller than this: k_collide walks the pairs
    int has_cyl;                             // a scene has cylinders (host side: the CYL kernels)
    int stale_ft, flat_inertia, keep_velocity, tmax_persist;
    int zero_force;  // k_finish leaves force/moment zero (the reference's state after a step)
};

// Queue segment and chunk of block b of a kernel that walks a queue with nblocks blocks (a
// multiple of kQueueSegs): segment b mod 64, chunk b / 64 of it.
struct QueueBlock {
    uint32_t seg, chunk, nchunk;
};

__device__ __forceinline__ QueueBlock queue_block(uint32_t b, uint32_t nblocks)
{
    return {b % kQueueSegs, b / kQueueSegs, nblocks / kQueueSegs};
}

__device__ __forceinline__ uint32_t scene_index(const StepArgs& A, uint32_t i, bool multi)
{
    return multi ? A.scene_of[i] : 0u;
}

// Work-group ids are dealt round-robin to the 8 XCDs (each with its own L2). This maps the
// id so that each XCD gets one contiguous eighth of the range: neighbouring slots, whose
// candidates overlap, then share an L2. The map is a permutation of [0, n).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n)
{
    constexpr uint32_t kXcd = 8;
    const uint32_t per = n / kXcd, extra = n % kXcd;   // the first `extra` XCDs get per + 1
    const uint32_t x = b % kXcd, r = b / kXcd;
    return x * per + min(x, extra) + r;
}

// Pair search (steps that rebuild the lists): for the sphere in sorted slot j, the ids of
// every neighbour within r_i + r_j + reach + skin, where reach bounds the JKR separation
// distance of the scene's pairs: every pair that can pass the reference's predicate
// (cdist > 0, or inside the JKR separation distance of a cohesive pair) before any sphere
// has moved half the skin. Buckets are at least 2 r_max + reach + skin wide, so the 27
// buckets around the sphere's hold them all. The three buckets of an x-row are consecutive
// keys, or two runs when the row wraps (the buckets up to the end of the row, then those
// from its start). One flat loop walks all nine rows, so each lane advances through its own
// candidates without waiting for the other lanes row by row. Only an FP32 test runs here;
// every step k_touch (whose build-step branch calls this) applies the exact FP64 predicate to
// the lists and puts a list in the reference's visiting order when its passing pairs are not
// in that order. Returns the sphere's list length and candidate flags; writes the list and
// lowers the sphere's gap to the least FP32 distance a listed pair lacks to pass.
//
// FP32 test: a candidate is dropped only when its FP32 distance exceeds r_i + r_j + reach +
// skin by more than the rounding of the FP32 coordinates (|x| 2^-24 per coordinate and
// operand, bounded by mg) and of the FP32 radius sum (factor 1 + 1e-6).
// The traversal works on bucket coordinates, i.e. the bucket-grid cell coordinates masked to
// the grid, exactly as grid_bucket(grid_cell(position)).
template <bool MULTI>
__device__ __forceinline__ uint32_t pairs_slot(const StepArgs& A, uint32_t j, uint32_t i)
{
    uint32_t n = 0;
    float pgap = 3.0e38f;
    const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
    const float4 fi = A.sposf[j];
    if (S.has_pp) {
        // the sphere's bucket coordinates from its key (dims are powers of two)
        const Grid& G = S.grid;
        const uint32_t kb = A.keys[i] - G.base;
        const uint32_t bx = kb & G.mask_x;
        const uint32_t by = (kb >> __builtin_ctz(G.dim_x)) & G.mask_y;
        const uint32_t bz = kb >> __builtin_ctz(G.dim_xy);
        const float mg = 2.4e-7f * (fabsf(fi.x) + fabsf(fi.y) + fabsf(fi.z)) + 1e-9f;
        // lim = r_j (1 + 1e-6) + [(r_i + reach)(1 + 1e-6) + mg]: the same bound with the
        // sphere's own part hoisted (each FP32 rounding is far below the 1e-6 margin)
        const float limi = ((fabsf(fi.w) + (float)(S.cc.coh_reach + S.skin * A.lctl->scale)) * 1.000001f) + mg;
        const float limr = ((fabsf(fi.w) + (float)S.cc.coh_reach) * 1.000001f) + 2.0f * mg + 1e-9f;
        const uint32_t lc = A.list_cap;
        uint32_t* out = A.nbr + (size_t)i * lc;
        const uint32_t bx0 = (bx - 1u) & G.mask_x;
        const uint32_t n1 = min(3u, G.dim_x - bx0);   // buckets before the wrap
        uint32_t row = 0, t = 0, s1 = 0, len1 = 0, s2 = 0, total = 0;
        auto load_row = [&](uint32_t r) {
            const uint32_t rb = G.base + ((bz + r / 3u - 1u) & G.mask_z) * G.dim_xy +
                                ((by + r % 3u - 1u) & G.mask_y) * G.dim_x;
            s1 = A.bucket_start[rb + bx0];
            const uint32_t e1 = A.bucket_start[rb + bx0 + n1];
            s2 = A.bucket_start[rb];
            const uint32_t e2 = A.bucket_start[rb + 3 - n1];
            len1 = e1 - s1;
            total = len1 + (e2 - s2);
            t = 0;
        };
        // candidate m with FP32 copy f, in visiting order; pgap: the least FP32 distance
        // a listed pair lacks to pass, less the rounding bound
        auto test = [&](uint32_t m, float4 f) {
            if (m == j) return;
            const float dx = f.x - fi.x, dy = f.y - fi.y, dz = f.z - fi.z;
            const float rj = fabsf(f.w) * 1.000001f;
            const float lim = rj + limi;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > lim * lim) return;
            pgap = fminf(pgap, sqrtf(d2) - (rj + limr));
            if (n < lc) out[n] = A.svals[m];
            n++;
        };
        load_row(0);
        while (true) {
            if (t < total) {
                // two candidates per iteration (both loads in flight), tested in order
                const uint32_t m0 = t < len1 ? s1 + t : s2 + (t - len1);
                const bool two = t + 1u < total;
                const uint32_t t1 = t + 1u;
                const uint32_t m1 = two ? (t1 < len1 ? s1 + t1 : s2 + (t1 - len1)) : m0;
                t += two ? 2u : 1u;
                const float4 f0 = A.sposf[m0];
                const float4 f1 = A.sposf[m1];
                test(m0, f0);
                if (two) test(m1, f1);
            } else {
                if (++row == 9u) break;
                load_row(row);
            }
        }
    }
    // candidates: spheres with listed neighbours or a wall in reach + skin (the order phase
    // marks the latter with the sign of the FP32 radius); k_touch tests them every step
    const bool wall_near = (__float_as_uint(fi.w) >> 31) != 0u;
    const bool full = n != 0u || wall_near;
    if (n) A.gap[i] = fminf(A.gap[i], pgap);
    return n | (full ? kCandidate : 0u) | (wall_near ? kWallNear : 0u);
}

// Triangle search (steps that rebuild the lists): every triangle of the scene's mesh whose
// closest point (the reference's closestPtPointTriangle, FP64) is within r + reach + skin of
// sphere i, from the body-frame grid cell of the sphere's centre (the grid lists every
// triangle within r_max + reach + the largest skin of any point of a cell). A triangle the
// reference's predicate can accept before either the sphere or the body has moved half the
// skin is within that distance now. Writes the sphere's triangle list and returns its length.
__device__ __forceinline__ uint32_t mesh_slot(const StepArgs& A, const SceneDev& S, uint32_t i)
{
    const MeshDev& M = A.meshes[S.mesh];
    const V4<double> p = ld4(A.pos + 4 * (size_t)i);
    const V4<double> e = {M.dbi[XDEM_E0], M.dbi[XDEM_E1], M.dbi[XDEM_E2], M.dbi[XDEM_E3]};
    const V3<double> loc = rot_local(mk3(p.x - M.dbi[XDEM_PX], p.y - M.dbi[XDEM_PY], p.z - M.dbi[XDEM_PZ]), e);
    const double fx = floor((loc.x - M.lo_x) / M.h), fy = floor((loc.y - M.lo_y) / M.h),
                 fz = floor((loc.z - M.lo_z) / M.h);
    uint32_t n = 0;
    if (fx >= 0.0 && fx < (double)M.dim_x && fy >= 0.0 && fy < (double)M.dim_y && fz >= 0.0 && fz < (double)M.dim_z) {
        const uint32_t c = M.cell_base + ((uint32_t)fz * M.dim_y + (uint32_t)fy) * M.dim_x + (uint32_t)fx;
        const uint32_t b = A.mcell_start[c], en = A.mcell_start[c + 1];
        const double lim = (p.w + M.cc.coh_reach + S.skin * (double)A.lctl->scale) * (1.0 + 1e-9) + 1e-12;
        const V3<double> ip = mk3(p.x, p.y, p.z);
        uint32_t* out = A.tri_list + (size_t)i * kTriCap;
        for (uint32_t k = b; k < en; k++) {
            const uint32_t t = A.mcell_tris[k];
            int ct;
            const V3<double> d = ip - closest_on_triangle(A.tri_glob[M.tri_begin + t], ip, ct);
            if (!(dot(d, d) <= lim * lim)) continue;
            if (n < kTriCap) out[n] = t;
            n++;
        }
        if (n > kTriCap) atomicOr(A.err, XDEM_ERR_MESH_NEAR);
    }
    const uint32_t m = min(n, kTriCap);
    A.tri_count[i] = m;
    return m;
}

// Drift of the next step (impl:135-147; xIntegratorVV.cpp:66-100): Euler parameters and
// position from the state the velocity update just produced, written to the second
// position/orientation buffer, which the next step makes current; and the half-skin test
// that decides whether the next step rebuilds the grid and the neighbour lists. Same
// expressions, in the same order, as a separate position-update pass would evaluate.
__device__ __forceinline__ V4<double> next_orientation(const SceneDev& S, V4<double> e, V4<double> d, V4<double> a)
{
    V4<double> ne;
    ne.x = e.x + S.dt * d.x + S.half2dt * a.x;
    ne.y = e.y + S.dt * d.y + S.half2dt * a.y;
    ne.z = e.z + S.dt * d.z + S.half2dt * a.z;
    ne.w = e.w + S.dt * d.w + S.half2dt * a.w;
    const double l = sqrt(ne.x * ne.x + ne.y * ne.y + ne.z * ne.z + ne.w * ne.w);
    ne.x = ne.x / l; ne.y = ne.y / l; ne.z = ne.z / l; ne.w = ne.w / l;
    return ne;
}

__device__ __forceinline__ void drift_orientation(const StepArgs& A, const SceneDev& S, uint32_t i, V4<double> e,
                                                  V4<double> d, V4<double> a)
{
    st4(A.ep_next + 4 * i, next_orientation(S, e, d, a));
}

__device__ __forceinline__ bool same_bits(V4<double> a, V4<double> b)
{
    return __double_as_longlong(a.x) == __double_as_longlong(b.x) && __double_as_longlong(a.y) == __double_as_longlong(b.y) &&
           __double_as_longlong(a.z) == __double_as_longlong(b.z) && __double_as_longlong(a.w) == __double_as_longlong(b.w);
}

__device__ __forceinline__ bool same_bits(V3<double> a, V3<double> b)
{
    return __double_as_longlong(a.x) == __double_as_longlong(b.x) && __double_as_longlong(a.y) == __double_as_longlong(b.y) &&
           __double_as_longlong(a.z) == __double_as_longlong(b.z);
}

__device__ __forceinline__ float drift_position(const StepArgs& A, const SceneDev& S, uint32_t i, V4<double> p,
                                                V3<double> v, V3<double> ac)
{
    const V4<double> p0 = p;
    p.x += S.dt * v.x + S.half2dt * ac.x;
    p.y += S.dt * v.y + S.half2dt * ac.y;
    p.z += S.dt * v.z + S.half2dt * ac.z;
    st4(A.pos_next + 4 * i, p);
    // half-skin test: the lists of the last build hold every pair within reach as long as no
    // sphere has moved half the skin since. The bound is the path length since the build
    // (disp, set to 0 by the build), each step's length rounded up; k_touch compares it with
    // the gaps of the lists.
    float dsp = 0.0f;
    if (A.rebuild_next) {
        const double dx = p.x - p0.x, dy = p.y - p0.y, dz = p.z - p0.z;
        dsp = __double2float_ru((double)A.disp[i] + sqrt(dx * dx + dy * dy + dz * dz) * (1.0 + 1e-6) + 1e-12);
        if (!(dsp < 0.5f * (float)(S.skin * A.lctl->scale))) {
            *A.rebuild_next = 1u;
            if (!(dsp < 3.0e38f)) dsp = 3.0e38f;   // a non-finite position: the largest bound
        }
        A.disp[i] = dsp;
    }
    return dsp;
}

__device__ __forceinline__ float drift_sphere(const StepArgs& A, const SceneDev& S, uint32_t i, V4<double> p,
                                              V4<double> e, V3<double> v, V3<double> ac, V4<double> d, V4<double> a)
{
    drift_orientation(A, S, i, e, d, a);
    return drift_position(A, S, i, p, v, ac);
}

// Wave-wide max of the displacement bounds into the segment counter of the wave (floats >= 0
// compare like their bits); every lane of the wave calls it.
__device__ __forceinline__ void add_disp_max(const StepArgs& A, float d)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) d = fmaxf(d, __shfl_xor(d, off, 64));
    if ((threadIdx.x & 63) == 0 && d > 0.0f && A.dmax_next) {
        const uint32_t w = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) & (kQueueSegs - 1u);
        atomicMax(&A.dmax_next[w * kSegStride], __float_as_uint(d));
    }
}

// A sphere's own state, loaded together before any branch (independent loads in flight at
// once). Fused mode also needs what the velocity update reads.
struct OwnState {
    V4<double> pos, ep, ev;
    V3<double> vel, t0;          // t0: carried rolling torque sum (zero unless it persists)
    double mass;
    V4<double> ea;               // fused mode only
    V3<double> acc, inertia;     // fused mode only
    double J;                    // fused mode only: rolling inertia (quirk Q5 aware)
};

// The fields of the velocity update (fused mode).
__device__ __forceinline__ void load_own_fused(const StepArgs& A, const SceneDev& S, uint32_t i, OwnState& o)
{
    o.ea = ld4(A.ea + 4 * i);
    o.acc = ld3(A.acc + 3 * i);
    o.inertia = ld3(A.inertia + 3 * i);
    // quirk Q5: the scene's inertia array read as flat doubles at the scene-local id (in the
    // caller's order), taken at commit
    o.J = A.flat_inertia ? A.jflat[i] : o.inertia.x;
}

// FUSED_TOO: also load the velocity-update fields now (the lite pass, which keeps few
// values live; the full pass loads them just before its tail instead).
template <bool FUSED_TOO>
__device__ __forceinline__ OwnState load_own(const StepArgs& A, const SceneDev& S, uint32_t i)
{
    OwnState o;
    o.pos = ld4(A.pos + 4 * i);
    o.vel = ld3(A.vel + 3 * i);
    o.ep = ld4(A.ep + 4 * i);
    o.ev = ld4(A.ev + 4 * i);
    o.mass = A.mass[i];
    o.t0 = A.tmax_persist ? ld3(A.tmax + 3 * i) : mk3(0.0, 0.0, 0.0);
    if (FUSED_TOO) load_own_fused(A, S, i, o);
    return o;
}

// ---------------------------------------------------------------- reference visiting order
// The reference visits the neighbours of sphere i cell by cell, z, y, x offsets -1..1 in
// that nesting, spheres of a cell by id (impl:613-733 over the sorted cell table; the stable
// sort of (hash, index), def:94-124). A pair that passes its predicate lies in neighbouring
// cells, so its place in that order is the key (code << 27 | id) with code =
// (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1) of the reference's cells (impl:54-72 without the
// 128 wrap, which only aliases cells too far apart to pass); code 27 marks a sphere whose
// cell is not a neighbour (the reference never visits it). Sphere ids stay below 2^27.
constexpr uint32_t kIdBits = 27;
constexpr uint32_t kIdMask = (1u << kIdBits) - 1u;
constexpr uint32_t kNotVisited = 27u;

__device__ __forceinline__ void ref_cell(const SceneDev& S, V4<double> p, int& cx, int& cy, int& cz)
{
    cx = (int)floor((p.x - S.grid.wo_x) / S.ref_cs);
    cy = (int)floor((p.y - S.grid.wo_y) / S.ref_cs);
    cz = (int)floor((p.z - S.grid.wo_z) / S.ref_cs);
}

__device__ __forceinline__ uint32_t visit_code(const SceneDev& S, int ix, int iy, int iz, V4<double> pj)
{
    int jx, jy, jz;
    ref_cell(S, pj, jx, jy, jz);
    const int dx = jx - ix, dy = jy - iy, dz = jz - iz;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return kNotVisited;
    return (uint32_t)((dz + 1) * 9 + (dy + 1) * 3 + (dx + 1));
}

// Puts a neighbour list in the reference's visiting order at the current positions: codes
// are packed above the caller's ids (the reference's ties within a cell), the list is
// insertion-sorted in place (it is short and this thread's own), then each entry is turned
// back into the stored sphere. k_touch does it when the pairs of a list that
// pass are not in that order (after a build, the lists are in bucket order; later, a sphere
// may have changed cells).
__device__ __forceinline__ void order_list(const StepArgs& A, const SceneDev& S, uint32_t* lst, uint32_t n, int ix,
                                           int iy, int iz)
{
    for (uint32_t t = 0; t < n; t++) {
        const uint32_t k = lst[t];
        lst[t] = (visit_code(S, ix, iy, iz, ld4(A.pos + 4 * (size_t)k)) << kIdBits) | A.ext_id[k];
    }
    for (uint32_t t = 1; t < n; t++) {
        const uint32_t x = lst[t];
        uint32_t u = t;
        while (u > 0u && lst[u - 1u] > x) {
            lst[u] = lst[u - 1u];
            u--;
        }
        lst[u] = x;
    }
    for (uint32_t t = 0; t < n; t++) lst[t] = A.int_id[lst[t] & kIdMask];
}

// The reference's pair predicate: overlap, or inside the JKR separation depth of a cohesive
// pair (impl:671-718).
__device__ __forceinline__ bool pair_passes(const ContactConst& cc, double ir, double jr, const PairGeom& g)
{
    if (g.cdist > 0) return true;
    if (!(cc.coh_reach > 0.0 && -g.cdist < cc.coh_reach)) return false;
    const double coh_s = cohesion_limit_depth(cc, jr ? (ir * jr) / (ir + jr) : ir);
    return fabs(g.cdist) < fabs(coh_s);
}

// Exact test of candidate i with list flags fl: walls in reach, the listed pairs that pass
// (in the reference's visiting order, the list re-sorted when broken), their mask. built: the
// lists were made this step (no displacement since); dm: the largest displacement since the
// build.
template <bool MULTI, bool CYL>
__device__ __forceinline__ bool touch_sphere(const StepArgs& A, uint32_t i, uint32_t fl, bool built, float dm)
{
    // a sphere with listed triangles goes to the full pass every step (the triangle pass
    // evaluates its list exactly)
    const bool mesh_near = (fl & kMeshNear) != 0u;
    bool touch = mesh_near;
    // a pair can pass only once the two spheres have moved its gap together, a wall come
    // into reach only once the sphere has moved its gap
    const float di = built ? 0.0f : A.disp[i];
    const bool may = mesh_near || !(A.gap[i] > (di + dm) * 1.0001f + 1e-9f);
    const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
    const V4<double> p = ld4(A.pos + 4 * i);
    if (may && (fl & kWallNear)) {
        for (uint32_t w = 0; w < S.nwall; w++) {
            const WallDev& W = A.walls[S.wall_begin + w];
            touch = touch || !wall_out_of_reach(W.cc, W.info, p);
        }
        if (CYL)
            for (uint32_t c = 0; c < S.ncyl; c++) {
                const CylDev& C = A.cyls[S.cyl_begin + c];
                touch = touch || !cylinder_out_of_reach(C.cc, C.info, C.dbi, C.margin, p, 0.0);
            }
    }
    const uint32_t n = may ? min(fl & kCountMask, A.list_cap) : 0u;
    if (n) {
        int ix, iy, iz;
        ref_cell(S, p, ix, iy, iz);
        uint32_t* lst = A.nbr + (size_t)i * A.list_cap;
        uint32_t prev = 0;
        bool have = false, ordered = true;
        unsigned long long m = 0;
        // four list entries at a time: their ids, then their positions, are loaded together
        for (uint32_t t0 = 0; t0 < n; t0 += 4u) {
            uint32_t kk[4];
            V4<double> pp[4];
#pragma unroll
            for (uint32_t u = 0; u < 4u; u++) kk[u] = t0 + u < n ? lst[t0 + u] : i;
#pragma unroll
            for (uint32_t u = 0; u < 4u; u++) pp[u] = ld4(A.pos + 4 * (size_t)kk[u]);
#pragma unroll
            for (uint32_t u = 0; u < 4u; u++) {
                const uint32_t t = t0 + u;
                if (t >= n) break;
                const V4<double> pj = pp[u];
                if (!pair_passes(S.cc, p.w, pj.w, pair_geom(p, pj))) continue;
                const uint32_t code = visit_code(S, ix, iy, iz, pj);
                if (code == kNotVisited) continue;
                m |= 1ull << t;
                const uint32_t key = (code << kIdBits) | A.ext_id[kk[u]];
                if (have && key <= prev) ordered = false;
                prev = key;
                have = true;
            }
        }
        if (!ordered) {
            // sort, then find the passing entries at their new places
            order_list(A, S, lst, n, ix, iy, iz);
            m = 0;
            for (uint32_t t = 0; t < n; t++) {
                const V4<double> pj = ld4(A.pos + 4 * (size_t)lst[t]);
                if (pair_passes(S.cc, p.w, pj.w, pair_geom(p, pj)) && visit_code(S, ix, iy, iz, pj) != kNotVisited)
                    m |= 1ull << t;
            }
        }
        touch = touch || m != 0ull;
        A.pass_mask[i] = m;
    }
    return touch;
}

// Appends the lanes with `on` set to queue segment seg (one atomic per wave, by its first
// active lane).
__device__ __forceinline__ void queue_append(uint32_t* q, uint32_t* count, uint32_t seg, uint32_t cap, bool on,
                                             uint32_t i)
{
    const unsigned long long mask = __ballot(on);
    if (!mask) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t leader = (uint32_t)__ffsll((long long)mask) - 1u;
    uint32_t b = 0;
    if (lane == leader) b = atomicAdd(&count[seg * kSegStride], (uint32_t)__popcll(mask));
    b = __shfl(b, (int)leader, 64);
    if (on) q[(size_t)seg * cap + b + __popcll(mask & ((1ull << lane) - 1ull))] = i;
}

// Every step, the spheres of this step's full pass. On a build step (the rebuild kernel's
// phases completed), one thread per sorted slot: the pair search for the slot's sphere
// (pairs_slot: its list, list gap and candidate flags, the candidate queue work), then the
// exact test of a candidate on the list it has just written. On the other steps, the
// candidates of the last build from their queue (block b walks segment b mod kQueueSegs):
// candidates whose gap exceeds their displacement plus the largest displacement since the
// build cannot touch and are not tested. Touched spheres get kTouch, the pass mask, and a
// place in the full-pass queue work2 (the segment of their candidate queue entry, so a
// segment never holds more than seg_cap); the others are completed by the lite pass, as the
// full pass would add only zero sums for them. The grid has a multiple of kQueueSegs
// blocks.
template <bool MULTI, bool CYL>
__global__ __launch_bounds__(256) void k_touch(StepArgs A)
{
    const bool built = *A.rebuild_cur != 0u;
    // A build always completes: a rebuild block that stops waiting (XDEM_ERR_GRID_SYNC, a
    // warning) has done every unit it claimed, and the block that completes a phase's last
    // unit passes its own wait and goes on, so every later phase still runs to the end. The
    // flags are checked anyway: without the last phase the bucket table may be inconsistent,
    // so then no list is read at all; every sphere loses kTouch and the lite pass completes
    // the step without contacts, and the error word says so.
    if (built && A.grid_bar[kSegStride * bar_flag(0)] < kRebuildPhases) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(A.err, XDEM_ERR_GRID_SYNC | XDEM_ERR_BUILD_INCOMPLETE);
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A.np; i += gridDim.x * blockDim.x)
            A.nbr_count[i] &= ~kTouch;
        return;
    }
    if (built) {
        // virtual blocks of 256 sorted slots, block vb on the XCD of physical block vb mod
        // gridDim (gridDim is a multiple of 8)
        const uint32_t nvb = (A.nsorted + blockDim.x - 1) / blockDim.x;
        for (uint32_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
            const uint32_t j = xcd_block(vb, nvb) * blockDim.x + threadIdx.x;
            const uint32_t i = j < A.nsorted ? A.svals[j] : kInvalid;
            bool full = false, touch = false, mesh = false;
            if (i < A.np) {
                uint32_t fl = pairs_slot<MULTI>(A, j, i);
                if (A.has_mesh) {
                    const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
                    if (S.mesh >= 0 && mesh_slot(A, S, i)) fl |= kMeshNear | kCandidate;
                    mesh = (fl & kMeshNear) != 0u;
                }
                full = (fl & kCandidate) != 0u;
                if (full) touch = touch_sphere<MULTI, CYL>(A, i, fl, true, 0.0f);
                A.nbr_count[i] = touch ? (fl | kTouch) : fl;
            }
            const uint32_t seg = (j >> 6) & (kQueueSegs - 1u);
            queue_append(A.work, A.seg_count, seg, A.seg_cap, full, i);
            queue_append(A.work2, A.seg_count2, seg, A.seg_cap, touch, i);
            if (A.has_mesh) queue_append(A.mesh_work, A.mesh_count, 0u, 0u, mesh, i);
        }
        return;
    }
    const QueueBlock qb = queue_block(blockIdx.x, gridDim.x);
    const uint32_t seg = qb.seg;
    const uint32_t nchunk = qb.nchunk;
    const uint32_t* wq = A.work + (size_t)seg * A.seg_cap;
    uint32_t base = qb.chunk * blockDim.x;
    // the first entry is read together with the segment's count
    uint32_t i_next = base + threadIdx.x < A.seg_cap ? wq[base + threadIdx.x] : 0u;
    const uint32_t cnt = A.seg_count[seg * kSegStride];
    if (base >= cnt) return;
    // no sphere has moved more than dm since the build (0 on a build step)
    __shared__ float dm_sh;
    if (threadIdx.x < 64) {
        float d = built ? 0.0f : __uint_as_float(A.dmax_cur[threadIdx.x * kSegStride]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) d = fmaxf(d, __shfl_xor(d, off, 64));
        if (threadIdx.x == 0) dm_sh = d;
    }
    __syncthreads();
    // moving walls and meshes: their displacement since the build bounds the wall part of
    // the gap the same way (taking the larger of the two keeps both bounds)
    const float dm = fmaxf(dm_sh, __double2float_ru(A.lctl->body_path - A.lctl->body_path_build));
    for (; base < cnt; base += nchunk * blockDim.x) {
        const uint32_t e = base + threadIdx.x;
        bool touch = false;
        const uint32_t i = i_next;
        if (e + nchunk * blockDim.x < cnt) i_next = wq[e + nchunk * blockDim.x];
        if (e < cnt) {
            const uint32_t fl = A.nbr_count[i];
            touch = touch_sphere<MULTI, CYL>(A, i, fl, built, dm);
            // a touched sphere's rotation, acceleration and contact counts are written by the
            // full pass: no longer known stable
            A.nbr_count[i] = touch ? ((fl & ~kLiteStable) | kTouch) : (fl & ~kTouch);
        }
        queue_append(A.work2, A.seg_count2, seg, A.seg_cap, touch, i);
    }
}

// End of a sphere's contact pass. Fused mode: rolling resistance (impl:2420-2454) and the
// velocity and orientation-rate update (xIntegratorVV.cpp:131-152; impl:192-222). Split
// mode: the sums are stored for k_finish.
template <bool FUSE>
__device__ __forceinline__ void sphere_tail(const StepArgs& A, const SceneDev& S, uint32_t i, const OwnState& o,
                                            Accum tot, uint32_t& err, float& dmx)
{
    if (FUSE) {
        const V4<double> e = o.ep;
        V4<double> av = o.ev;
        const bool consumed = rolling_moment(S.dt, o.J, omega_of(e, av), tot.T, tot.R, tot.M);
        if (A.tmax_persist) st3(A.tmax + 3 * i, consumed ? mk3(0.0, 0.0, 0.0) : tot.T);
        const double m = o.mass;
        const V4<double> aa = o.ea;
        const V3<double> np_ = rot_local(tot.M, e);
        const V4<double> mea = uceom(o.inertia, e, av, np_);
        const double inv_m = 1.0 / m;
        const V3<double> a = mk3(inv_m * (tot.F.x + m * S.gx), inv_m * (tot.F.y + m * S.gy),
                                 inv_m * (tot.F.z + m * S.gz));
        const double hdt = 0.5 * S.dt;
        V3<double> v = o.vel;
        if (A.keep_velocity) {
            av.x = av.x + hdt * (aa.x + mea.x); av.y = av.y + hdt * (aa.y + mea.y);
            av.z = av.z + hdt * (aa.z + mea.z); av.w = av.w + hdt * (aa.w + mea.w);
            if (A.vel_out != A.vel) st3(A.vel_out + 3 * i, v);
        } else {
            const V3<double> ao = o.acc;
            v = mk3(v.x + hdt * ao.x, v.y + hdt * ao.y, v.z + hdt * ao.z);
            av.x = av.x + hdt * aa.x; av.y = av.y + hdt * aa.y; av.z = av.z + hdt * aa.z; av.w = av.w + hdt * aa.w;
            v = mk3(v.x + hdt * a.x, v.y + hdt * a.y, v.z + hdt * a.z);
            av.x = av.x + hdt * mea.x; av.y = av.y + hdt * mea.y;
            av.z = av.z + hdt * mea.z; av.w = av.w + hdt * mea.w;
            st3(A.vel_out + 3 * i, v);
        }
        st4(A.ev_out + 4 * i, av);
        st3(A.acc + 3 * i, a);
        st4(A.ea + 4 * i, mea);
        if (!(isfinite(a.x) && isfinite(a.y) && isfinite(a.z))) err |= XDEM_ERR_NONFINITE;
        if (A.drift) dmx = fmaxf(dmx, drift_sphere(A, S, i, o.pos, e, v, a, av, mea));
    } else {
        st3(A.force + 3 * i, tot.F);
        st3(A.moment + 3 * i, tot.M);
        st3(A.tmax + 3 * i, tot.T);
        A.rres[i] = tot.R;
    }
}

// ------------------------------------------------------------------------- pair pass
// The sphere contacts of the full pass, one thread per queued sphere: its passing pairs in
// rank order (rank = place among the sphere's passing list entries, i.e. the reference's
// visiting order), each looked up in the previous step's slots, evaluated
// (sphere_pair_eval), written to its new slot (= rank, up to the capacity) and added with
// sphere_contact's operations, so the sums are those of the reference's walk over the pairs.
// The first kPairIdRegs slot ids of the previous step are read once into registers, and the
// next pair's list entry is loaded while the current pair is evaluated; the neighbour's state
// and its history entry are then one round trip. The sums go to ppsum for k_collide.
// Cohesion-only pairs add their attraction.
constexpr uint32_t kPairIdRegs = 16;
constexpr uint32_t kPairSet = 0, kPairStale = 1, kPairCoh = 2;   // kinds of an evaluated pair

// The first slot of the previous step holding sphere k (kInvalid: none).
__device__ __forceinline__ uint32_t old_slot(const StepArgs& A, uint32_t hb, const uint32_t (&ids)[kPairIdRegs],
                                             uint32_t old_count, uint32_t k)
{
    uint32_t slot = kInvalid;
#pragma unroll
    for (int s = (int)kPairIdRegs - 1; s >= 0; s--)   // downwards: the lowest match stays
        if ((uint32_t)s < old_count && ids[s] == k) slot = (uint32_t)s;
    for (uint32_t s = kPairIdRegs; s < old_count && slot == kInvalid; s++)
        if (A.pp_id_in[hb + s] == k) slot = s;
    return slot;
}

// Sphere i's pair with neighbour k: the neighbour's state, the history entry of `slot`, the
// force law; writes the new slot `rank` and returns the kind.
template <class T, int MODEL, bool GPU_LAW>
__device__ __forceinline__ uint32_t pair_of(const StepArgs& A, const SceneDev& S, uint32_t i, const OwnState& o,
                                            V3<double> iom, uint32_t k, uint32_t rank, uint32_t slot, double v[10])
{
    const uint32_t hb = i * A.cap;
    const V4<double> jpos = ld4(A.pos + 4 * (size_t)k);
    const V3<double> jvel = ld3(A.vel + 3 * (size_t)k);
    const V4<double> jep = ld4(A.ep + 4 * (size_t)k);
    const V4<double> jev = ld4(A.ev + 4 * (size_t)k);
    const double jmass = A.mass[k];
    const double2 h = slot != kInvalid ? A.pp_tsd_in[hb + slot] : make_double2(0.0, 0.0);
    const PairGeom g = pair_geom(o.pos, jpos);
    double sd0 = 0.0, sd1 = 0.0;
    uint32_t kind;
    if (g.cdist > 0) {
        sd0 = h.x; sd1 = h.y;
        const bool set = sphere_pair_eval<T, MODEL, GPU_LAW>(S.cc, S.dt, o.pos, o.vel, iom, o.mass, jpos, jvel,
                                                             omega_of(jep, jev), jmass, g, sd0, sd1, v);
        kind = set ? kPairSet : kPairStale;
    } else {
        // cohesion only: the attraction cf * unit that sphere_cohesion subtracts
        const double ir = o.pos.w, jr = jpos.w;
        const double Req = jr ? (ir * jr) / (ir + jr) : ir;
        const double cf = cohesion_force<double>(S.cc, Req, S.cc.jk * Req);
        const V3<double> u = g.rp / g.dist;
        const V3<double> a = cf * u;
        v[0] = a.x; v[1] = a.y; v[2] = a.z;
        kind = kPairCoh;
    }
    if (rank < A.cap) {
        A.pp_id_out[hb + rank] = k;
        A.pp_tsd_out[hb + rank] = make_double2(sd0, sd1);
    }
    return kind;
}

// The full-pass spheres of this step (all queue segments): one counter per lane of a wave.
__device__ __forceinline__ uint32_t full_queue_total(const StepArgs& A)
{
    static_assert(kQueueSegs == 64, "one segment counter per lane");
    uint32_t c = A.seg_count2[(threadIdx.x & 63u) * kSegStride];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    return c;
}

// Below StepArgs::merge_below full-pass spheres (default kMergePairsBelow) k_collide walks
// the pairs itself and k_pairs returns at once: the spheres then fit in one wave per SIMD of
// k_collide, and the pair sums skip two dependent memory round trips (ppsum, own state).
// Above it k_pairs (three waves per SIMD) hides the neighbour gathers of dense packs.
constexpr uint32_t kMergePairsBelow = 65536;

// Sphere i's passing pairs in rank order (the k_pairs walk): the pair sums, the new slots and
// count, the overflow bits. Returns the stored count.
template <class T, int MODEL, bool GPU_LAW>
__device__ __forceinline__ uint32_t pairs_sphere(const StepArgs& A, const SceneDev& S, uint32_t i, const OwnState& o,
                                                 V3<double> iom, uint32_t flags, Accum& pp)
{
    const uint32_t nlist_all = flags & kCountMask;
    const unsigned long long m = min(nlist_all, A.list_cap) ? A.pass_mask[i] : 0ull;
    const uint32_t old_count = min(A.pp_count_in[i], A.cap);
    const uint32_t hb = i * A.cap;
    // the previous step's slot ids (at most the capacity: the last sphere's slots end the array)
    uint32_t ids[kPairIdRegs];
    const uint32_t nid = min(A.cap, kPairIdRegs);
    if ((A.cap & 3u) == 0u) {
        const uint4* p4 = reinterpret_cast<const uint4*>(A.pp_id_in + hb);
#pragma unroll
        for (uint32_t c = 0; c < kPairIdRegs / 4u; c++) {
            const uint4 w = 4u * c < nid ? p4[c] : make_uint4(kInvalid, kInvalid, kInvalid, kInvalid);
            ids[4 * c] = w.x; ids[4 * c + 1] = w.y; ids[4 * c + 2] = w.z; ids[4 * c + 3] = w.w;
        }
    } else {
#pragma unroll
        for (uint32_t s = 0; s < kPairIdRegs; s++) ids[s] = s < nid ? A.pp_id_in[hb + s] : kInvalid;
    }
    const uint32_t* lst = A.nbr + (size_t)i * A.list_cap;
    V3<T> Ft = mk3(T(0), T(0), T(0));
    uint32_t rank = 0;
    unsigned long long mm = m;
    uint32_t k_next = mm ? lst[(uint32_t)__ffsll((long long)mm) - 1u] : 0u;
    for (; mm; rank++) {
        const uint32_t k = k_next;
        mm &= mm - 1ull;
        if (mm) k_next = lst[(uint32_t)__ffsll((long long)mm) - 1u];   // in flight during this pair
        double v[10];
        const uint32_t kind =
            pair_of<T, MODEL, GPU_LAW>(A, S, i, o, iom, k, rank, old_slot(A, hb, ids, old_count, k), v);
        if (kind == kPairCoh) pp.F = pp.F - mk3(v[0], v[1], v[2]);
        else sphere_pair_sum<T>(v, kind == kPairSet, A.stale_ft != 0, Ft, pp);
    }
    const uint32_t n_new = min(rank, A.cap);
    A.pp_count_out[i] = n_new;
    uint32_t err = 0;
    if (rank > A.cap) err |= XDEM_ERR_P2P_OVERFLOW;
    if (nlist_all > A.list_cap) err |= XDEM_ERR_NBR_OVERFLOW;
    if (err) atomicOr(A.err, err);
    return n_new;
}

// Sphere i against the cylinders of its scene, in order: each looks up the slots as the previous
// one left them (one kernel launch per cylinder in the reference); the slots are updated in
// place and the stored count returned. The full-pass and grid kernels take it only in their
// CYL instantiation (contexts with cylinders): its registers stay out of the others.
template <class T, int MODEL, bool GPU_LAW>
__device__ __forceinline__ uint32_t collide_cylinders(const CylDev* cyls, uint32_t ncyl, uint32_t* cy_count,
                                                   uint32_t* cy_id, double2* cy_tsd, double* wrench, uint32_t np,
                                                   uint32_t i, double dt, V4<double> pos, V3<double> vel,
                                                   V3<double> iom, double mass, Accum& tot, uint32_t& err)
{
    CylSlots cs;
    const uint32_t cb = XDEM_MAX_P2CY_COUNT * i;
#pragma unroll
    for (uint32_t q = 0; q < XDEM_MAX_P2CY_COUNT; q++) {
        cs.id[q] = cy_id[cb + q];
        const double2 h = cy_tsd[cb + q];
        cs.s0[q] = h.x;
        cs.s1[q] = h.y;
    }
    cs.count = cy_count[i];
    cs.overflow = false;
    for (uint32_t c = 0; c < ncyl; c++) {
        const CylDev& C = cyls[c];
#pragma unroll
        for (uint32_t q = 0; q < XDEM_MAX_P2CY_COUNT; q++) {
            cs.old_id[q] = cs.id[q]; cs.old_s0[q] = cs.s0[q]; cs.old_s1[q] = cs.s1[q];
        }
        double bf[6] = {0, 0, 0, 0, 0, 0};
        cylinder_contact<T, MODEL, GPU_LAW>(C.cc, C.info, C.dbi, dt, pos, vel, iom, mass, cs, tot, bf);
        if (C.wslot >= 0) {
            double* wr = wrench + (size_t)C.wslot * 6 * np;
#pragma unroll
            for (int q = 0; q < 6; q++) wr[(size_t)q * np + i] = bf[q];
        }
    }
    if (cs.overflow) err |= XDEM_ERR_CYL_OVERFLOW;
#pragma unroll
    for (uint32_t q = 0; q < XDEM_MAX_P2CY_COUNT; q++) {
        cy_id[cb + q] = cs.id[q];
        cy_tsd[cb + q] = make_double2(cs.s0[q], cs.s1[q]);
    }
    cy_count[i] = cs.count;
    return cs.count;
}

// Full contact pass of sphere i: the near neighbours the pair search listed, in their order, under
// the exact FP64 predicate; every wall of the scene in id order; then sphere_tail. The
// sphere's own state and outputs are read and written by id.
template <class T, bool MULTI, bool FUSE, int MODEL, bool GPU_LAW, bool CYL>
__device__ __forceinline__ void collide_sphere(const StepArgs& A, uint32_t i, bool merged, uint32_t& n_pp,
                                               uint32_t& n_wall, float& dmx)
{
    const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
    // the sphere's own state, list flags and wall slot count, issued together (the fields
    // of the velocity update are read just before the tail: fewer registers live across
    // the pairs)
    OwnState o = load_own<false>(A, S, i);
    const uint32_t flags = A.nbr_count[i];
    n_pp = 0;
    WallSlots ws;
    const uint32_t count0 = S.nwall ? A.pl_count[i] : 0u;
    ws.init(count0, A.pl_id + XDEM_MAX_P2PL_COUNT * i,
            reinterpret_cast<const double*>(A.pl_tsd + XDEM_MAX_P2PL_COUNT * i));
    const V3<double> iom = omega_of(o.ep, o.ev);
    uint32_t err = 0;

    // ---- sphere-sphere (impl:613-733): walked here when the full pass is small, else the
    // sums of k_pairs, which ran before this kernel
    Accum pp = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), 0.0};
    if (S.has_pp) {
        if (merged) {
            n_pp = pairs_sphere<T, MODEL, GPU_LAW>(A, S, i, o, iom, flags, pp);
        } else {
            const double* ps = A.ppsum + (size_t)i * 10;
            pp.F = mk3(ps[0], ps[1], ps[2]);
            pp.M = mk3(ps[3], ps[4], ps[5]);
            pp.T = mk3(ps[6], ps[7], ps[8]);
            pp.R = ps[9];
        }
    }

    // ---- walls, in id order (impl:950-1080 once per wall)
    Accum tot;
    tot.F = mk3(0.0, 0.0, 0.0) + pp.F;
    tot.M = mk3(0.0, 0.0, 0.0) + pp.M;
    tot.T = o.t0 + pp.T;
    tot.R = 0.0 + pp.R;
    if (S.nwall && !(flags & kWallNear)) {
        // no wall in reach + skin since the list build (the build's exact test): every wall's
        // contact is empty, so the slot count follows wall_no_contact's rule as in the lite
        // pass, and a dynamic wall's wrench slot gets zeros
        const uint32_t count = S.wall_id0 ? 0u : min(count0, (uint32_t)XDEM_MAX_P2PL_COUNT);
        if (count != count0) A.pl_count[i] = count;
        n_wall = count;
        if (S.wall_dynamic)
            for (uint32_t w = 0; w < S.nwall; w++) {
                const int32_t slot = A.walls[S.wall_begin + w].wslot;
                if (slot < 0) continue;
                double* wr = A.wrench + (size_t)slot * 6 * A.np;
#pragma unroll
                for (int c = 0; c < 6; c++) wr[(size_t)c * A.np + i] = 0.0;
            }
    } else if (S.nwall) {
        // wall slots are read only when a wall needs them and written back only
        // when one changed (the values in memory are the same either way)
        for (uint32_t w = 0; w < S.nwall; w++) {
            const WallDev& W = A.walls[S.wall_begin + w];
            double bf[6] = {0, 0, 0, 0, 0, 0};
            if (wall_contact<T, MODEL, GPU_LAW>(W.cc, W.info, W.dbi, S.dt, o.pos, o.vel, iom, o.mass, ws, tot, bf))
                err |= XDEM_ERR_PLANE_OVERFLOW;
            if (W.wslot >= 0) {
                double* wr = A.wrench + (size_t)W.wslot * 6 * A.np;
#pragma unroll
                for (int c = 0; c < 6; c++) wr[(size_t)c * A.np + i] = bf[c];
            }
        }
        if (ws.dirty) {
#pragma unroll
            for (int s = 0; s < XDEM_MAX_P2PL_COUNT; s++) {
                A.pl_id[XDEM_MAX_P2PL_COUNT * i + s] = ws.id[s];
                A.pl_tsd[XDEM_MAX_P2PL_COUNT * i + s] = make_double2(ws.s0[s], ws.s1[s]);
            }
        }
        if (ws.count != count0) A.pl_count[i] = ws.count;
        n_wall = ws.count;
    }

    // ---- cylinders, in the scene's order (cylinder_contact_force_kernel once per cylinder,
    // impl:1327-1460)
    if (CYL && S.ncyl)
        n_wall += collide_cylinders<T, MODEL, GPU_LAW>(A.cyls + S.cyl_begin, S.ncyl, A.cy_count, A.cy_id, A.cy_tsd,
                                                       A.wrench, A.np, i, S.dt, o.pos, o.vel, iom, o.mass, tot, err);

    // ---- the scene's mesh: this step's triangle sums (k_mesh_pass), added after the walls, as
    // the reference's mesh launch adds its sums to force, moment, tmax and rres (impl:2100-2108)
    if (A.has_mesh && S.mesh >= 0) {
        if (flags & kMeshNear) {
            const double* ms = A.msum + (size_t)i * 10;
            tot.F = tot.F + mk3(ms[0], ms[1], ms[2]);
            tot.M = tot.M + mk3(ms[3], ms[4], ms[5]);
            tot.T = tot.T + mk3(ms[6], ms[7], ms[8]);
            tot.R += ms[9];
        } else {
            tot.F = tot.F + mk3(0.0, 0.0, 0.0);
            tot.M = tot.M + mk3(0.0, 0.0, 0.0);
            tot.T = tot.T + mk3(0.0, 0.0, 0.0);
            tot.R += 0.0;
            A.ms_count[i] = 0u;   // no triangle within reach: the reference's mesh pass stores none
        }
    }

    if (FUSE) load_own_fused(A, S, i, o);
    sphere_tail<FUSE>(A, S, i, o, tot, err, dmx);
    if (err) atomicOr(A.err, err);
}

// ----------------------------------------------------------------- triangle pass
// Triangle contacts of sphere i with its scene's mesh (impl:1930-2109 for one sphere): the
// listed triangles are taken in the reference's visiting order (cells of the proxy sphere
// centres, z, y, x offsets, then triangle index: the key (code << 27) | t as for spheres);
// contacts (and cohesive near-contacts) are sorted into faces, edges and vertices, an edge or
// vertex point equal within 1e-9 to the previous one of its kind is dropped (the previous
// points start at the origin, as in the reference), and faces apply, else edges, else
// vertices. Writes the sphere's triangle sums (k_collide adds them after its walls), its
// body-wrench contribution and its triangle history (5 slots shared by all meshes, in place:
// the old slots are read into registers first, as the reference's kernel does).
template <class T, bool MULTI, int MODEL, bool GPU_LAW>
__device__ __forceinline__ void mesh_sphere(const StepArgs& A, uint32_t i)
{
    constexpr uint32_t CAP = XDEM_MAX_P2MS_COUNT;
    const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
    const MeshDev& M = A.meshes[S.mesh];
    const V4<double> p4 = ld4(A.pos + 4 * (size_t)i);
    const V3<double> ip = mk3(p4.x, p4.y, p4.z);
    const double ir = p4.w;
    const V3<double> ivel = ld3(A.vel + 3 * (size_t)i);
    const V3<double> iom = omega_of(ld4(A.ep + 4 * (size_t)i), ld4(A.ev + 4 * (size_t)i));
    const double im = A.mass[i];
    uint32_t pid[CAP];
    double ps0[CAP], ps1[CAP];
#pragma unroll
    for (uint32_t q = 0; q < CAP; q++) {
        pid[q] = A.ms_id[CAP * (size_t)i + q];
        const double2 h = A.ms_tsd[CAP * (size_t)i + q];
        ps0[q] = h.x;
        ps1[q] = h.y;
    }
    double dbi[15];
#pragma unroll
    for (int q = 0; q < 15; q++) dbi[q] = M.dbi[q];
    const ContactConst& cc = M.cc;
    const double coh_s = cc.coh ? cohesion_limit_depth(cc, ir) : 0.0;
    // the listed triangles that touch (or are inside the cohesion depth), by visiting key
    int ix, iy, iz;
    ref_cell(S, p4, ix, iy, iz);
    const uint32_t n = min(A.tri_count[i], kTriCap);
    const uint32_t* lst = A.tri_list + (size_t)i * kTriCap;
    uint32_t key[kTriCap];
    uint32_t nk = 0;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t t = lst[k];
        int ct;
        const V3<double> cpt = closest_on_triangle(A.tri_glob[M.tri_begin + t], ip, ct);
        const double cdist = ir - length(ip - cpt);
        if (!(cdist > 0 || fabs(cdist) < fabs(coh_s))) continue;
        const double* pc = A.prox_glob + 4 * (size_t)(M.tri_begin + t);
        const uint32_t code = visit_code(S, ix, iy, iz, V4<double>{pc[0], pc[1], pc[2], 0.0});
        if (code == kNotVisited) continue;
        uint32_t x = (code << kIdBits) | t, u = nk;
        while (u > 0u && key[u - 1u] > x) {   // insertion sort
            key[u] = key[u - 1u];
            u--;
        }
        key[u] = x;
        nk++;
    }
    // the reference's selection in visiting order: which of the sorted entries are faces,
    // edges and vertices
    uint32_t faces = 0, edges = 0, verts = 0;
    V3<double> prev_line = mk3(0.0, 0.0, 0.0), prev_point = mk3(0.0, 0.0, 0.0);
    for (uint32_t k = 0; k < nk; k++) {
        const uint32_t t = key[k] & kIdMask;
        int ct;
        const V3<double> cpt = closest_on_triangle(A.tri_glob[M.tri_begin + t], ip, ct);
        const double cdist = ir - length(ip - cpt);
        if (cdist > 0) {
            if (ct == 1) {
                if (same_point(prev_line, cpt)) continue;
                prev_line = cpt;
            } else if (ct == 0) {
                if (same_point(prev_point, cpt)) continue;
                prev_point = cpt;
            }
        }
        if (ct == 2) faces |= 1u << k;
        else if (ct == 1) edges |= 1u << k;
        else if (ct == 0) verts |= 1u << k;
    }
    const uint32_t apply = faces ? faces : (edges ? edges : verts);
    Accum acc = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), 0.0};
    double bf[6] = {0, 0, 0, 0, 0, 0};
    uint32_t n_new = 0, n_seen = 0;
    for (uint32_t m = apply; m; m &= m - 1u) {
        const uint32_t k = (uint32_t)__ffs((int)m) - 1u;
        const uint32_t t = key[k] & kIdMask;
        int ct;
        const TriHit h = {t, closest_on_triangle(A.tri_glob[M.tri_begin + t], ip, ct)};
        double sd0, sd1;
        if (triangle_contact<T, MODEL, GPU_LAW>(cc, dbi, S.dt, h, ip, ir, im, ivel, iom, pid, ps0, ps1, sd0, sd1, acc,
                                                bf)) {
            n_seen++;
            if (n_new < CAP) {
                A.ms_id[CAP * (size_t)i + n_new] = t;
                A.ms_tsd[CAP * (size_t)i + n_new] = make_double2(sd0, sd1);
                n_new++;
            }
        }
    }
    A.ms_count[i] = n_new;
    double* ms = A.msum + (size_t)i * 10;
    st3(ms, acc.F);
    st3(ms + 3, acc.M);
    st3(ms + 6, acc.T);
    ms[9] = acc.R;
    if (M.wslot >= 0) {
        double* wr = A.wrench + (size_t)M.wslot * 6 * A.np;
#pragma unroll
        for (int c = 0; c < 6; c++) wr[(size_t)c * A.np + i] = bf[c];
    }
    if (n_seen > CAP) atomicOr(A.err, XDEM_ERR_MESH_OVERFLOW);
}

// The triangle pass over the spheres with listed triangles (the mesh queue of the last build);
// it runs before k_collide, which adds the sums.
template <class T, bool MULTI, int MODEL, bool GPU_LAW>
__global__ __launch_bounds__(256) void k_mesh_pass(StepArgs A)
{
    const uint32_t cnt = *A.mesh_count;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += gridDim.x * blockDim.x) {
        const uint32_t i = A.mesh_work[e];
        if (i < A.np) mesh_sphere<T, MULTI, MODEL, GPU_LAW>(A, i);
    }
}

// Pair results handed from the lane that evaluated a pair to the lane of its sphere, per wave:
// ten values (sphere_pair_eval's out[] or the cohesion attraction) and the kind, column-major.
struct PairStage {
    double v[10][64];
    uint32_t kind[64];
    uint32_t ids[64][kPairIdRegs];   // each lane's sphere: its first slot ids of the previous step
};

// LDS written by some lanes of a wave and then read by others: the wave's own accesses complete
// in program order; the fences keep the compiler from moving them across this point.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The pair pass of a large full pass. A wave takes 64 touched spheres, one per lane: when at
// least half the sorted slots are touched (a dense pack), found in sorted-slot order, each XCD
// sweeping one contiguous eighth of the slots, so the neighbours a block gathers were mostly
// fetched into its XCD's L2 by the blocks of adjacent slots; otherwise from the full-pass
// queue (block b takes segment b mod 64). It numbers their passing pairs consecutively, sphere after sphere, in rank order (an exclusive
// scan of the pass-mask counts). The lanes then evaluate the pairs 64 at a time: entry x
// belongs to the lane L whose range [s_L, s_L + c_L) holds it (a binary search over the
// scanned ends), its rank is x - s_L and its list entry the rank-th set bit of L's mask; the
// pair's history lookup, force law and new slot are pair_of's, as in pairs_sphere. Each
// sphere's lane then adds its pairs' results in rank order from LDS with sphere_pair_sum's
// operations, so the sums, slots and counts are those of the one-thread walk, while no lane
// idles through the longest walk of its wave.
template <class T, bool MULTI, int MODEL, bool GPU_LAW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_pairs(StepArgs A)
{
    const uint32_t nfull = full_queue_total(A);
    if (nfull < A.merge_below) return;   // k_collide walks the pairs
    __shared__ PairStage stage_all[4];
    PairStage& st = stage_all[threadIdx.x >> 6];
    const uint32_t lane = threadIdx.x & 63u;
    unsigned long long c0 = 0;
    // the wave's spheres: sorted slots (it = virtual block) or queue entries (it = the wave's
    // first entry); either way the loop bounds are the same for the whole wave
    const bool sweep = 2ull * nfull >= (unsigned long long)A.nsorted;
    const uint32_t nvb = (A.nsorted + blockDim.x - 1) / blockDim.x;
    const QueueBlock qb = queue_block(blockIdx.x, gridDim.x);
    const uint32_t* qseg = A.work2 + (size_t)qb.seg * A.seg_cap;
    const uint32_t qcnt = sweep ? 0u : A.seg_count2[qb.seg * kSegStride];
    const uint32_t it0 = sweep ? blockIdx.x : qb.chunk * blockDim.x + (threadIdx.x & ~63u);
    const uint32_t it_end = sweep ? nvb : qcnt;
    const uint32_t it_step = sweep ? gridDim.x : qb.nchunk * blockDim.x;
    for (uint32_t it = it0; it < it_end; it += it_step) {
        uint32_t i = kInvalid;
        if (sweep) {
            const uint32_t j = xcd_block(it, nvb) * blockDim.x + threadIdx.x;
            if (j < A.nsorted) i = A.svals[j];
        } else if (it + lane < qcnt) {
            i = qseg[it + lane];
        }
        uint32_t flags = 0;
        if (i < A.np) {
            flags = A.nbr_count[i];
            if (!(flags & kTouch) || !A.scenes[scene_index(A, i, MULTI)].has_pp) i = kInvalid;
        } else {
            i = kInvalid;
        }
        const uint32_t nlist_all = flags & kCountMask;
        const unsigned long long m = (i != kInvalid && min(nlist_all, A.list_cap)) ? A.pass_mask[i] : 0ull;
        const uint32_t c = (uint32_t)__popcll(m);
        const uint32_t old_count = c ? min(A.pp_count_in[i], A.cap) : 0u;
        if (c) {
            const uint32_t hb = i * A.cap;
            const uint32_t nid = min(old_count, kPairIdRegs);
            for (uint32_t q = 0; q < nid; q++) st.ids[lane][q] = A.pp_id_in[hb + q];
        }
        // inclusive scan of the counts over the wave
        uint32_t incl = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off, 64);
            if ((int)lane >= off) incl += t;
        }
        const uint32_t total = __shfl(incl, 63, 64);
        const uint32_t first = incl - c;
        Accum pp = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), 0.0};
        V3<T> Ft = mk3(T(0), T(0), T(0));
        wave_lds_sync();
        // entry x's lane L, rank r and neighbour k; the next chunk's are found (and its list
        // entry loaded) while this chunk's pairs are evaluated
        uint32_t L, r, k;
        auto locate = [&](uint32_t x, uint32_t& L_, uint32_t& r_, uint32_t& k_) {
            // the lane of entry x: the number of lanes whose range ends at or before x
            L_ = 0;
#pragma unroll
            for (uint32_t w = 32; w; w >>= 1) {
                const uint32_t e = __shfl(incl, (int)(L_ + w - 1u), 64);
                if (e <= x) L_ += w;
            }
            const uint32_t iq = __shfl(i, (int)L_, 64);
            r_ = x - __shfl(first, (int)L_, 64);
            const uint32_t mlo = __shfl((uint32_t)m, (int)L_, 64), mhi = __shfl((uint32_t)(m >> 32), (int)L_, 64);
            k_ = 0;
            if (x < total) {
                unsigned long long mm = ((unsigned long long)mhi << 32) | mlo;
                for (uint32_t u = 0; u < r_; u++) mm &= mm - 1ull;
                k_ = A.nbr[(size_t)iq * A.list_cap + ((uint32_t)__ffsll((long long)mm) - 1u)];
            }
        };
        locate(lane, L, r, k);
        for (uint32_t b = 0; b < total; b += 64u) {
            const uint32_t x = b + lane;
            uint32_t Ln = 0, rn = 0, kn = 0;
            if (b + 64u < total) locate(x + 64u, Ln, rn, kn);
            const uint32_t iL = __shfl(i, (int)L, 64);
            const uint32_t ocL = __shfl(old_count, (int)L, 64);
            if (x < total) {
                const SceneDev& S = A.scenes[scene_index(A, iL, MULTI)];
                OwnState o;
                o.pos = ld4(A.pos + 4 * (size_t)iL);
                o.vel = ld3(A.vel + 3 * (size_t)iL);
                o.mass = A.mass[iL];
                const V3<double> iom = omega_of(ld4(A.ep + 4 * (size_t)iL), ld4(A.ev + 4 * (size_t)iL));
                const uint32_t hb = iL * A.cap;
                uint32_t slot = kInvalid;
#pragma unroll
                for (int q = (int)kPairIdRegs - 1; q >= 0; q--)   // downwards: the lowest match stays
                    if ((uint32_t)q < ocL && st.ids[L][q] == k) slot = (uint32_t)q;
                for (uint32_t q = kPairIdRegs; q < ocL && slot == kInvalid; q++)
                    if (A.pp_id_in[hb + q] == k) slot = q;
                double v[10];
                const uint32_t kind = pair_of<T, MODEL, GPU_LAW>(A, S, iL, o, iom, k, r, slot, v);
#pragma unroll
                for (int q = 0; q < 10; q++) st.v[q][lane] = v[q];
                st.kind[lane] = kind;
            }
            wave_lds_sync();
            // this lane's sphere adds its pairs of this chunk, in rank order
            const uint32_t lo = max(first, b), hi = min(incl, b + 64u);
            for (uint32_t y = lo; y < hi; y++) {
                const uint32_t q = y - b;
                double v[10];
#pragma unroll
                for (int z = 0; z < 10; z++) v[z] = st.v[z][q];
                const uint32_t kind = st.kind[q];
                if (kind == kPairCoh) pp.F = pp.F - mk3(v[0], v[1], v[2]);
                else sphere_pair_sum<T>(v, kind == kPairSet, A.stale_ft != 0, Ft, pp);
            }
            wave_lds_sync();
            L = Ln; r = rn; k = kn;
        }
        if (i != kInvalid) {
            const uint32_t n_new = min(c, A.cap);
            A.pp_count_out[i] = n_new;
            c0 += n_new;
            uint32_t err = 0;
            if (c > A.cap) err |= XDEM_ERR_P2P_OVERFLOW;
            if (nlist_all > A.list_cap) err |= XDEM_ERR_NBR_OVERFLOW;
            if (err) atomicOr(A.err, err);
            double* ps = A.ppsum + (size_t)i * 10;
            st3(ps, pp.F);
            st3(ps + 3, pp.M);
            st3(ps + 6, pp.T);
            ps[9] = pp.R;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c0 += __shfl_xor(c0, off, 64);
    if ((threadIdx.x & 63) == 0 && c0) atomicAdd(&A.counters[0], c0);
}

// Wave-wide sums of the stored-contact counts, one atomic per wave.
__device__ __forceinline__ void add_contact_counts(const StepArgs& A, unsigned long long c0, unsigned long long c1)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        c0 += __shfl_xor(c0, off, 64);
        c1 += __shfl_xor(c1, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && (c0 | c1)) {
        atomicAdd(&A.counters[0], c0);
        atomicAdd(&A.counters[1], c1);
    }
}

// Lite pass: every sphere outside this step's full pass (no listed pair passes, no wall in
// reach): the full pass would add only zero sums, so the lite pass applies the same tail to
// zero sums, with the walls' slot-count rule of wall_no_contact taken from the scene's
// summary (the count drops to 0 if the scene has a wall with id 0, impl:989). All loads are
// issued before the first branch. It runs concurrently with the full pass (the two touch
// disjoint spheres).

// What the lite pass reads of one sphere, all issued together.
struct LiteIn {
    OwnState o;
    uint32_t flags;    // nbr_count (kTouch: the full pass owns the sphere this step)
    uint32_t count0;   // wall slot count
};

template <bool FUSE>
__device__ __forceinline__ void lite_load(const StepArgs& A, const SceneDev& S, uint32_t i, LiteIn& in)
{
    in.o = load_own<FUSE>(A, S, i);
    in.flags = A.nbr_count[i];
    in.count0 = S.nwall ? A.pl_count[i] : 0u;
}

template <bool FUSE>
__device__ __forceinline__ void lite_sphere(const StepArgs& A, const SceneDev& S, uint32_t i, const LiteIn& in,
                                            uint32_t skip, uint32_t& n_wall)
{
    if (in.flags & skip) return;
    uint32_t err = 0;
    if (S.has_pp) A.pp_count_out[i] = 0u;
    Accum tot;
    tot.F = mk3(0.0, 0.0, 0.0);
    tot.M = mk3(0.0, 0.0, 0.0);
    tot.T = in.o.t0 + mk3(0.0, 0.0, 0.0);
    tot.R = 0.0;
    if (S.mesh >= 0) A.ms_count[i] = 0u;
    if (S.nwall) {
        const uint32_t count = S.wall_id0 ? 0u : min(in.count0, (uint32_t)XDEM_MAX_P2PL_COUNT);
        if (count != in.count0) A.pl_count[i] = count;
        n_wall += count;
    }
    if (S.ncyl) {
        const uint32_t c0 = A.cy_count[i];
        const uint32_t count = S.cyl_id0 ? 0u : min(c0, (uint32_t)XDEM_MAX_P2CY_COUNT);
        if (count != c0) A.cy_count[i] = count;
        n_wall += count;
    }
    float dmx = 0.0f;   // the split lite pass computes no drift
    sphere_tail<FUSE>(A, S, i, in.o, tot, err, dmx);
    if (err) atomicOr(A.err, err);
}

// k_rebuild's phase cursors, counters and release flags start at zero: cleared by block 0 of
// the lite pass (after k_touch has read the release flags, before the next step's
// k_rebuild).
__device__ __forceinline__ void reset_rebuild_slots(const StepArgs& A)
{
    if (blockIdx.x == 0)
        for (uint32_t t = threadIdx.x; t < kBarSlots; t += blockDim.x) A.grid_bar[kSegStride * t] = 0u;
}

// The lite kernels run over every sphere outside this step's full pass (skip kTouch), one
// per thread.
template <class F>
__device__ __forceinline__ void lite_spheres(const StepArgs& A, F&& body)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.np) body(i, kTouch);
}

// The lite pass of a split step (contact sums only).
template <class T, bool MULTI, bool FUSE, int MODEL, bool GPU_LAW>
__global__ __launch_bounds__(256) void k_collide_lite(StepArgs A)
{
    reset_rebuild_slots(A);
    uint32_t n_wall = 0;
    lite_spheres(A, [&](uint32_t i, uint32_t skip) {
        const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
        LiteIn in;
        lite_load<FUSE>(A, S, i, in);
        lite_sphere<FUSE>(A, S, i, in, skip, n_wall);
    });
    add_contact_counts(A, 0u, n_wall);
}

// The lite pass of a fused step, split by motion: k_lite_move does the translation of every
// lite sphere (acceleration, velocity, next position) and its contact bookkeeping,
// k_lite_turn the rotation (carried rolling torque, Euler-parameter rates, next orientation).
// They read and write disjoint fields, so each needs few registers and runs at a high
// occupancy. Together they evaluate the expressions of sphere_tail with zero contact sums,
// in the same order.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_lite_move(StepArgs A)
{
    reset_rebuild_slots(A);
    uint32_t n_wall = 0;
    float dmx = 0.0f;
    lite_spheres(A, [&](uint32_t i, uint32_t skip) {
        const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
        const uint32_t flags = A.nbr_count[i];
        const bool settled = flags & kNoContacts;
        const bool acc_kept = flags & kAccStable;
        const uint32_t count0 = (S.nwall && !settled) ? A.pl_count[i] : 0u;
        const uint32_t pp_in = (S.has_pp && !settled) ? A.pp_count_in[i] : 0u;
        const uint32_t msc = (S.mesh >= 0 && !settled) ? A.ms_count[i] : 0u;
        const uint32_t cyc0 = (S.ncyl && !settled) ? A.cy_count[i] : 0u;
        const V4<double> p = ld4(A.pos + 4 * i);
        const V3<double> v0 = ld3(A.vel + 3 * i);
        const V3<double> ao = ld3(A.acc + 3 * i);
        const double m = acc_kept ? 0.0 : A.mass[i];
        if (flags & skip) return;
        uint32_t nf = flags;
        if (!settled) {
            // contact bookkeeping of a sphere without contacts: empty p2p and triangle history,
            // the walls' and cylinders' slot-count rule; once all of it holds zeros (the p2p
            // count in both buffers) nothing changes until the sphere is touched again
            if (S.has_pp) A.pp_count_out[i] = 0u;
            if (msc) A.ms_count[i] = 0u;
            uint32_t count = 0u, cyc = 0u;
            if (S.nwall) {
                count = S.wall_id0 ? 0u : min(count0, (uint32_t)XDEM_MAX_P2PL_COUNT);
                if (count != count0) A.pl_count[i] = count;
                n_wall += count;
            }
            if (S.ncyl) {
                cyc = S.cyl_id0 ? 0u : min(cyc0, (uint32_t)XDEM_MAX_P2CY_COUNT);
                if (cyc != cyc0) A.cy_count[i] = cyc;
                n_wall += cyc;
            }
            if (pp_in == 0u && count == 0u && msc == 0u && cyc == 0u) nf |= kNoContacts;
        }
        V3<double> a = ao;
        if (!acc_kept) {
            const double inv_m = 1.0 / m;
            a = mk3(inv_m * (0.0 + m * S.gx), inv_m * (0.0 + m * S.gy), inv_m * (0.0 + m * S.gz));
            const bool finite = isfinite(a.x) && isfinite(a.y) && isfinite(a.z);
            if (!finite) atomicOr(A.err, XDEM_ERR_NONFINITE);
            if (!same_bits(a, ao)) st3(A.acc + 3 * i, a);
            else if (finite) nf |= kAccStable;   // the next steps compute the same bits
        }
        const double hdt = 0.5 * S.dt;
        V3<double> v = v0;
        if (A.keep_velocity) {
            if (A.vel_out != A.vel) st3(A.vel_out + 3 * i, v);
        } else {
            v = mk3(v.x + hdt * ao.x, v.y + hdt * ao.y, v.z + hdt * ao.z);
            v = mk3(v.x + hdt * a.x, v.y + hdt * a.y, v.z + hdt * a.z);
            st3(A.vel_out + 3 * i, v);
        }
        if (nf != flags) A.nbr_count[i] = nf;
        if (A.drift) dmx = fmaxf(dmx, drift_position(A, S, i, p, v, a));
    });
    add_contact_counts(A, 0u, n_wall);
    add_disp_max(A, dmx);
}

// A lite sphere's rotation update depends only on ep, ev, ea, inertia and the carried
// rolling torque. When it leaves them all unchanged (no spin, no torque) it is at a fixed
// point: kSpinStable is set once both buffers of ev and ep hold those bits, and from then on
// the sphere is skipped (nothing read, nothing written) until k_touch or a list build clears
// the flag.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_lite_turn(StepArgs A)
{
    lite_spheres(A, [&](uint32_t i, uint32_t skip) {
        const uint32_t flags = A.nbr_count[i];
        if (flags & (skip | kSpinStable)) return;
        const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
        const V4<double> e = ld4(A.ep + 4 * i);
        const V4<double> ev0 = ld4(A.ev + 4 * i);
        const V4<double> aa = ld4(A.ea + 4 * i);
        const V3<double> inertia = ld3(A.inertia + 3 * i);
        const V3<double> t0 = A.tmax_persist ? ld3(A.tmax + 3 * i) : mk3(0.0, 0.0, 0.0);
        // no contact: the rolling sums are zero, so rolling_moment leaves the moment and the
        // carried torque as they are
        V3<double> M = mk3(0.0, 0.0, 0.0);
        const V3<double> T = t0 + mk3(0.0, 0.0, 0.0);
        if (A.tmax_persist) st3(A.tmax + 3 * i, T);
        V4<double> av = ev0;
        const V3<double> np_ = rot_local(M, e);
        const V4<double> mea = uceom(inertia, e, av, np_);
        const double hdt = 0.5 * S.dt;
        if (A.keep_velocity) {
            av.x = av.x + hdt * (aa.x + mea.x); av.y = av.y + hdt * (aa.y + mea.y);
            av.z = av.z + hdt * (aa.z + mea.z); av.w = av.w + hdt * (aa.w + mea.w);
        } else {
            av.x = av.x + hdt * aa.x; av.y = av.y + hdt * aa.y; av.z = av.z + hdt * aa.z; av.w = av.w + hdt * aa.w;
            av.x = av.x + hdt * mea.x; av.y = av.y + hdt * mea.y;
            av.z = av.z + hdt * mea.z; av.w = av.w + hdt * mea.w;
        }
        const bool ea_same = same_bits(mea, aa);
        if (!ea_same) st4(A.ea + 4 * i, mea);
        st4(A.ev_out + 4 * i, av);
        if (!A.drift) return;
        const V4<double> ne = next_orientation(S, e, av, mea);
        st4(A.ep_next + 4 * i, ne);
        const bool t_same = !A.tmax_persist || (__double_as_longlong(T.x) == __double_as_longlong(t0.x) &&
                                                __double_as_longlong(T.y) == __double_as_longlong(t0.y) &&
                                                __double_as_longlong(T.z) == __double_as_longlong(t0.z));
        if (ea_same && t_same && same_bits(av, ev0) && same_bits(ne, e)) A.nbr_count[i] = flags | kSpinStable;
    });
}

// Full pass over the spheres k_touch queued: block b takes queue segment b mod kQueueSegs,
// its entries chunk b / kQueueSegs, ... strided by the chunks per segment. The first entry's
// sphere id is read together with the segment's count.
template <class T, bool MULTI, bool FUSE, int MODEL, bool GPU_LAW, bool CYL>
__global__ __launch_bounds__(256) void k_collide(StepArgs A)
{
    const QueueBlock qb = queue_block(blockIdx.x, gridDim.x);
    const uint32_t stride = qb.nchunk * blockDim.x;
    const uint32_t* q = A.work2 + (size_t)qb.seg * A.seg_cap;
    uint32_t e = qb.chunk * blockDim.x + threadIdx.x;
    uint32_t i = e < A.seg_cap ? q[e] : 0u;
    const uint32_t cnt = A.seg_count2[qb.seg * kSegStride];
    unsigned long long c0 = 0, c1 = 0;
    float dmx = 0.0f;
    const bool merged = full_queue_total(A) < A.merge_below;
    while (e < cnt) {
        uint32_t n_pp = 0, n_wall = 0;
        collide_sphere<T, MULTI, FUSE, MODEL, GPU_LAW, CYL>(A, i, merged, n_pp, n_wall, dmx);
        c0 += n_pp;
        c1 += n_wall;
        e += stride;
        if (e < cnt) i = q[e];
    }
    add_contact_counts(A, c0, c1);
    add_disp_max(A, dmx);
}

// Second half of a split step: rolling resistance and the velocity update from the
// force/moment/tmax/rres arrays (xContactManager.cpp:471, xIntegratorVV.cpp:62).
template <bool MULTI>
__global__ __launch_bounds__(256) void k_finish(StepArgs A)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float dmx = 0.0f;
    if (i < A.np) {
        const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
        const V4<double> e = ld4(A.ep + 4 * i);
        V4<double> av = ld4(A.ev + 4 * i);
        V3<double> M = ld3(A.moment + 3 * i);
        const double J = A.flat_inertia ? A.jflat[i] : A.inertia[3 * i];
        if (rolling_moment(S.dt, J, omega_of(e, av), ld3(A.tmax + 3 * i), A.rres[i], M)) {
            st3(A.tmax + 3 * i, mk3(0.0, 0.0, 0.0));
            A.rres[i] = 0.0;
        }
        const double m = A.mass[i];
        const V3<double> F = ld3(A.force + 3 * i);
        const V4<double> aa = ld4(A.ea + 4 * i);
        const V3<double> np_ = rot_local(M, e);
        const V4<double> mea = uceom(ld3(A.inertia + 3 * i), e, av, np_);
        const double inv_m = 1.0 / m;
        const V3<double> a = mk3(inv_m * (F.x + m * S.gx), inv_m * (F.y + m * S.gy), inv_m * (F.z + m * S.gz));
        const double hdt = 0.5 * S.dt;
        if (A.keep_velocity) {
            av.x = av.x + hdt * (aa.x + mea.x); av.y = av.y + hdt * (aa.y + mea.y);
            av.z = av.z + hdt * (aa.z + mea.z); av.w = av.w + hdt * (aa.w + mea.w);
        } else {
            V3<double> v = ld3(A.vel + 3 * i);
            const V3<double> ao = ld3(A.acc + 3 * i);
            v = mk3(v.x + hdt * ao.x, v.y + hdt * ao.y, v.z + hdt * ao.z);
            av.x = av.x + hdt * aa.x; av.y = av.y + hdt * aa.y; av.z = av.z + hdt * aa.z; av.w = av.w + hdt * aa.w;
            v = mk3(v.x + hdt * a.x, v.y + hdt * a.y, v.z + hdt * a.z);
            av.x = av.x + hdt * mea.x; av.y = av.y + hdt * mea.y; av.z = av.z + hdt * mea.z; av.w = av.w + hdt * mea.w;
            st3(A.vel + 3 * i, v);
        }
        st4(A.ev + 4 * i, av);
        st3(A.acc + 3 * i, a);
        st4(A.ea + 4 * i, mea);
        if (A.zero_force) {
            st3(A.force + 3 * i, mk3(0.0, 0.0, 0.0));
            st3(A.moment + 3 * i, mk3(0.0, 0.0, 0.0));
        }
        if (!(isfinite(a.x) && isfinite(a.y) && isfinite(a.z))) atomicOr(A.err, XDEM_ERR_NONFINITE);
        if (A.drift) dmx = drift_sphere(A, S, i, ld4(A.pos + 4 * i), e, ld3(A.vel + 3 * i), a, av, mea);
    }
    add_disp_max(A, dmx);
}

// Contact kernels of one precision (xdem_step_f64.hip / xdem_step_mixed.hip); pass 0 is the
// lite pass, pass 1 the full pass, pass 2 the triangle pass. The other kernels are FP64 only.
using CollideLaunch = hipError_t (*)(const StepArgs&, bool multi, bool fuse, int model, bool gpu_law, int pass,
                                     hipStream_t);
// which: 0 = k_lite_move then k_lite_turn, 1 = k_lite_move only, 2 = k_lite_turn only
void launch_lite_fused(const StepArgs&, bool multi, hipStream_t, int which = 0);
// Blocks of k_touch and k_pairs: as many as the device holds at once at three waves per SIMD
// (both kernels fit three: k_pairs by its launch bounds, k_touch by its registers), one
// round with no tail, a multiple of the queue segments.
inline uint32_t resident_grid(int cus)
{
    const uint32_t b = 3u * (uint32_t)(cus > 0 ? cus : 1) / kQueueSegs * kQueueSegs;
    return b > kQueueSegs ? b : kQueueSegs;
}
CollideLaunch collide_launch_f64();
CollideLaunch collide_launch_mixed();
// xdem_grid.hip
uint32_t bucket_table_size(uint32_t nbucket);   // bucket_count / bucket_start entries (+1 for start)
hipError_t launch_drift(const StepArgs&, bool multi, hipStream_t);
hipError_t launch_grid(const StepArgs&, bool multi, hipStream_t);   // k_rebuild, every step
uint32_t rebuild_grid_size(bool multi);   // k_rebuild's grid on the current device (0: error)
hipError_t launch_touch(const StepArgs&, bool multi, hipStream_t);   // k_touch (with the pair search), every step
hipError_t launch_finish(const StepArgs&, bool multi, hipStream_t);

}  // namespace xdem

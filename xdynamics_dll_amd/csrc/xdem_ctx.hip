// xdem_ctx.hip -- the step layer of the C ABI: device state of one or many scenes and
// the fused step (position update + bucket key, bucket counting sort + sorted copies,
// contacts + Verlet), optionally recorded in a HIP graph.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "xdem_capi.h"
#include "xdem_sort.hpp"
#include "xdem_step.hpp"

using namespace xdem;

namespace {

struct SceneHost {
    xdem_scene s;               // deep copy of the caller's arrays
    std::vector<double> pos, vel, mass, inertia, dbi;
    std::vector<xdem_plane_info> planes;
    std::vector<xdem_contact_property> plane_cp;
    std::vector<double> wall_local;   // [12 * nplane] corners in the wall's body frame, fixed at
                                      // add_scene (xPlaneObject::define's local_point)
    uint32_t p_begin = 0;
    uint32_t wall_begin = 0;
    double rmax = 0.0;                // largest sphere radius
    SceneDev dev;
};

struct MeshHost {
    uint32_t scene = 0;
    std::vector<double> vlist;        // [9 T] vertices, body frame
    std::vector<double> plocal;       // [3 T] proxy sphere centres, body frame
    std::vector<double> prad;         // [T] proxy radii
    std::vector<uint32_t> cell_start; // body-frame grid (local cell indices, local triangle order)
    std::vector<uint32_t> cell_tris;
    double rmax_local = 0.0;          // largest distance of a vertex from the body origin
    MeshDev dev;
};

struct CylHost {
    uint32_t scene = 0;
    double extent = 0.0;              // body-frame radius of the geometry a sphere can reach
    CylDev dev;
};

// A dynamic body's wrench: the spheres of [begin, end) whose flags hold `mask` wrote their
// contribution to its per-sphere slot this step (kTouch for a wall: the full pass writes every
// dynamic wall's slot; kMeshNear for a mesh: the triangle pass writes its slot).
struct WrenchSlot {
    uint32_t begin, end, mask, pad;
};

// A body pose handed to the device by value (kernel argument): the host's array may change
// as soon as the launch returns, with no copy in flight and no wait.
struct PoseArg {
    double v[15];
};

__global__ void k_set_mesh_pose(MeshDev* meshes, uint32_t m, PoseArg p)
{
    const uint32_t q = threadIdx.x;
    if (q < 15) meshes[m].dbi[q] = p.v[q];
}

// Triangles and proxy centres of mesh m at its pose (impl:2544-2601; the proxies of the first
// mesh are the reference's proxy spheres 0..T-1, tid = index).
__global__ __launch_bounds__(256) void k_mesh_pose(const MeshDev* meshes, uint32_t m, const double* vlist,
                                                  const double* plocal, const double* prad,
                                                  xdem_triangle_info* tri, double* prox)
{
    const MeshDev& M = meshes[m];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M.ntri) return;
    const size_t g = M.tri_begin + t;
    const V3<double> bp = mk3(M.dbi[XDEM_PX], M.dbi[XDEM_PY], M.dbi[XDEM_PZ]);
    const V4<double> e = {M.dbi[XDEM_E0], M.dbi[XDEM_E1], M.dbi[XDEM_E2], M.dbi[XDEM_E3]};
    const double* v = vlist + 9 * g;
    const V3<double> P = bp + rot_global(ld3(v + 0), e);
    const V3<double> Q = bp + rot_global(ld3(v + 3), e);
    const V3<double> R = bp + rot_global(ld3(v + 6), e);
    const V3<double> c = bp + rot_global(ld3(plocal + 3 * g), e);
    xdem_triangle_info ti;
    ti.id = (int)t;
    ti.tid = t;
    ti.P = {P.x, P.y, P.z};
    ti.Q = {Q.x, Q.y, Q.z};
    ti.R = {R.x, R.y, R.z};
    tri[g] = ti;
    prox[4 * g + 0] = c.x; prox[4 * g + 1] = c.y; prox[4 * g + 2] = c.z; prox[4 * g + 3] = prad[g];
}

// Wrench of every dynamic body, in a fixed order (the same bits on every run): block (b, s)
// sums the flagged spheres of chunk b of slot s's range, each thread a strided subset, then
// the waves and the block in a fixed order; k_wrench_final adds the kWrenchBlocks partials of
// each slot the same way.
__global__ __launch_bounds__(256) void k_wrench_partial(const double* wrench, const uint32_t* flags, uint32_t np,
                                                       const WrenchSlot* slots, double* partial)
{
    __shared__ double sh[6][4];
    const uint32_t slot = blockIdx.y, b = blockIdx.x;
    const WrenchSlot W = slots[slot];
    const uint32_t n = W.end - W.begin;
    const uint32_t chunk = (n + kWrenchBlocks - 1) / kWrenchBlocks;
    const uint32_t lo = W.begin + min(n, b * chunk), hi = W.begin + min(n, (b + 1) * chunk);
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256)
        if (flags[i] & W.mask) {
#pragma unroll
            for (int c = 0; c < 6; c++) v[c] += wrench[((size_t)slot * 6 + c) * np + i];
        }
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const double x = wave_sum(v[c]);
        if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < 6)
        partial[((size_t)slot * kWrenchBlocks + b) * 6 + threadIdx.x] =
            ((sh[threadIdx.x][0] + sh[threadIdx.x][1]) + sh[threadIdx.x][2]) + sh[threadIdx.x][3];
}

__global__ __launch_bounds__(256) void k_wrench_final(const double* partial, double* out)
{
    __shared__ double sh[6][4];
    const uint32_t slot = blockIdx.x;
    double v[6];
#pragma unroll
    for (int c = 0; c < 6; c++) v[c] = partial[((size_t)slot * kWrenchBlocks + threadIdx.x) * 6 + c];
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const double x = wave_sum(v[c]);
        if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < 6)
        out[(size_t)slot * 6 + threadIdx.x] =
            ((sh[threadIdx.x][0] + sh[threadIdx.x][1]) + sh[threadIdx.x][2]) + sh[threadIdx.x][3];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Rows of w 32-bit words in a new order: dst row j = src row perm[j].
__global__ void k_gather_rows(const uint32_t* src, uint32_t* dst, const uint32_t* perm, size_t n, uint32_t w)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * w) return;
    const size_t j = t / w, c = t - j * w;
    dst[t] = src[(size_t)perm[j] * w + c];
}

// Columns of n doubles (structure-of-arrays) in a new order: dst[col][j] = src[col][perm[j]].
__global__ void k_gather_cols(const double* src, double* dst, const uint32_t* perm, size_t n, uint32_t ncol)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * ncol) return;
    const size_t col = t / n, j = t - col * n;
    dst[t] = src[col * n + perm[j]];
}

// Sphere-pair slot ids in a new order: row j from row perm[j], each stored id below the (new)
// count renamed by inv (old stored index -> new).
__global__ void k_gather_pair_ids(const uint32_t* src, uint32_t* dst, const uint32_t* count, const uint32_t* perm,
                                  const uint32_t* inv, size_t n, uint32_t cap)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * cap) return;
    const size_t j = t / cap, q = t - j * cap;
    const uint32_t v = src[(size_t)perm[j] * cap + q];
    dst[t] = q < count[j] ? inv[v] : v;
}

// Rotation matrix of Euler parameters e (xPointMass::setupTransformationMatrix,
// source/xdynamics_object/xPointMass.cpp:296-301): o = A v, or A^T v when transpose is set.
void euler_rotate(const double* e, const double v[3], bool transpose, double o[3])
{
    const double a00 = 2 * (e[0] * e[0] + e[1] * e[1] - 0.5), a01 = 2 * (e[1] * e[2] - e[0] * e[3]), a02 = 2 * (e[1] * e[3] + e[0] * e[2]);
    const double a10 = 2 * (e[1] * e[2] + e[0] * e[3]), a11 = 2 * (e[0] * e[0] + e[2] * e[2] - 0.5), a12 = 2 * (e[2] * e[3] - e[0] * e[1]);
    const double a20 = 2 * (e[1] * e[3] - e[0] * e[2]), a21 = 2 * (e[2] * e[3] + e[0] * e[1]), a22 = 2 * (e[0] * e[0] + e[3] * e[3] - 0.5);
    if (!transpose) {
        o[0] = a00 * v[0] + a01 * v[1] + a02 * v[2]; o[1] = a10 * v[0] + a11 * v[1] + a12 * v[2]; o[2] = a20 * v[0] + a21 * v[1] + a22 * v[2];
    } else {
        o[0] = a00 * v[0] + a10 * v[1] + a20 * v[2]; o[1] = a01 * v[0] + a11 * v[1] + a21 * v[2]; o[2] = a02 * v[0] + a12 * v[1] + a22 * v[2];
    }
}

// A wall or mesh body moved by at most d since the last pose: the lists of the last build
// stay complete while the summed motion stays below half the smallest skin; beyond, the
// next step rebuilds them.
__global__ void k_body_moved(ListCtl* L, double d, uint32_t* rebuild_next)
{
    L->body_path += d;
    if (!(L->body_path - L->body_path_build < L->half_skin_min * (double)L->scale)) *rebuild_next = 1u;
}

uint32_t next_pow2(uint32_t v)
{
    uint32_t p = 1;
    while (p < v && p < 0x80000000u) p <<= 1;
    return p;
}

}  // namespace

struct xdem_ctx {
    xdem_ctx_config cfg;
    std::vector<SceneHost> scenes;
    std::vector<WallDev> walls;
    std::vector<MeshHost> meshes;
    std::vector<CylHost> cyls;          // cylinder bodies in the order added
    std::vector<uint32_t> cyl_dev;      // cylinder -> its entry of d_cyls (grouped by scene)
    std::vector<WrenchSlot> wslots;     // per dynamic body (wall, cylinder or mesh): sphere range and flag
    std::vector<int> wall_slot;         // global wall -> slot or -1
    bool committed = false;
    bool multi = false;
    int model = -1;
    uint32_t np = 0, nbucket = 0, key_bits = 1, cap = XDEM_MAX_P2P_COUNT, nwslot = 0;
    int tmax_persist = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;                       // the full contact pass
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;     // pair search done / full pass done
    CollideLaunch collide_fn = nullptr;

    double *pos = nullptr, *vel = nullptr, *acc = nullptr, *ep = nullptr, *ev = nullptr, *ea = nullptr;
    double *force = nullptr, *moment = nullptr, *mass = nullptr, *inertia = nullptr, *tmax = nullptr, *rres = nullptr;
    uint32_t* scene_of = nullptr;
    SceneDev* d_scenes = nullptr;
    WallDev* d_walls = nullptr;
    CylDev* d_cyls = nullptr;
    uint32_t *cy_count = nullptr, *cy_id = nullptr;
    double2* cy_tsd = nullptr;
    uint32_t *keys = nullptr, *svals = nullptr, *slot_tmp = nullptr;
    uint32_t *nbr_count = nullptr, *nbr = nullptr, *work = nullptr, *seg_count = nullptr;
    uint32_t *work2 = nullptr, *seg_count2 = nullptr;
    uint32_t* grid_bar = nullptr;                        // k_rebuild's phase cursors and flags
    uint32_t rebuild_blocks = 0;                         // k_rebuild's grid
    uint32_t resident_blocks = kQueueSegs;               // k_touch's and k_pairs' grid
    unsigned long long* pass_mask = nullptr;
    // stored order: stored sphere -> the caller's global index (scene begin + caller's index)
    // and back; the same ranges per scene as the caller's
    std::vector<uint32_t> ext_of, int_of;
    uint32_t *d_ext = nullptr, *d_int = nullptr;
    double* d_jflat = nullptr;   // quirk Q5's rolling inertia per stored sphere
    uint32_t seg_cap = 0;
    uint32_t list_cap = 48;
    uint32_t *bucket_count = nullptr, *bucket_start = nullptr, *scan_part = nullptr;
    float4* sposf = nullptr;
    // list rebuild flags, one per step parity: step s reads rebuild[rbp] (rbp flips at the
    // start of every step) and its drift sets rebuild[rbp ^ 1] for step s + 1
    uint32_t* rebuild = nullptr;
    int rbp = 1;
    uint32_t* dmax = nullptr;          // two sets of kQueueSegs counters, by rbp like rebuild
    float *gap = nullptr, *disp = nullptr;
    unsigned long long* n_rebuild = nullptr;
    ListCtl* lctl = nullptr;
    // vel and ev: two buffers each in fused mode (the contact kernel writes the new values
    // into the other one while neighbours still read the old ones); vel/ev point at the
    // current one
    double* vel_buf[2] = {nullptr, nullptr};
    double* ev_buf[2] = {nullptr, nullptr};
    int sv = 0;
    // pos and ep: two buffers each; a step's velocity update writes the next step's drift
    // into the other one (drift_valid), which the next step makes current
    double* pos_buf[2] = {nullptr, nullptr};
    double* ep_buf[2] = {nullptr, nullptr};
    int spp = 0;
    bool drift_valid = false;
    uint32_t* pp_count[2] = {nullptr, nullptr};
    uint32_t* pp_id[2] = {nullptr, nullptr};
    double2* pp_tsd[2] = {nullptr, nullptr};
    int cur = 0;
    uint32_t* pl_count = nullptr;
    uint32_t* pl_id = nullptr;
    double2* pl_tsd = nullptr;
    double* wrench_pp = nullptr;
    double* wrench_out = nullptr;
    double* h_wrench = nullptr;   // pinned host copy of one body's wrench (the coupled loop's readback)
    double* wrench_partial = nullptr;
    WrenchSlot* d_wslots = nullptr;
    // meshes (all scenes' triangles concatenated)
    MeshDev* d_meshes = nullptr;
    double *d_vlist = nullptr, *d_plocal = nullptr, *d_prad = nullptr, *d_prox = nullptr, *msum = nullptr;
    double* ppsum = nullptr;
    xdem_triangle_info* d_tri = nullptr;
    uint32_t *d_mcell_start = nullptr, *d_mcell_tris = nullptr;
    uint32_t *tri_count = nullptr, *tri_list = nullptr, *mesh_work = nullptr, *mesh_count = nullptr;
    uint32_t *ms_count = nullptr, *ms_id = nullptr;
    double2* ms_tsd = nullptr;
    uint32_t ntri = 0;
    uint32_t* err = nullptr;
    unsigned long long* counters = nullptr;
    std::vector<void*> allocs;
    size_t bytes = 0;

    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    unsigned graph_steps = 0;
    // xdem_ctx_step's own graph of auto_steps steps (replayed when the parities match)
    hipGraph_t agraph = nullptr;
    hipGraphExec_t agexec = nullptr;
    // one-step graphs for runs shorter than auto_steps (a coupled loop steps one at a time): a
    // step flips every buffer parity, so one graph per position parity covers every step
    hipGraph_t sgraph[2] = {nullptr, nullptr};
    hipGraphExec_t sgexec[2] = {nullptr, nullptr};
    int spar[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    unsigned auto_steps = 0;
    int apar[4] = {0, 0, 0, 0};
    int graph_parity = 0, graph_sparity = 0, graph_pparity = 0;

    bool timing = false;
    std::vector<hipEvent_t> events;  // kEventsPerStep per timed step
    double t_contact = 0.0, t_grid = 0.0, t_finish = 0.0, t_pairs = 0.0, t_lite_move = 0.0, t_collide = 0.0;
    double t_pairpass = 0.0, t_meshpass = 0.0;
    unsigned t_steps = 0;
    unsigned long long steps_done = 0;

    template <class P> hipError_t alloc(P** p, size_t bytes_)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes_ ? bytes_ : 16);
        if (e != hipSuccess) return e;
        allocs.push_back(q);
        bytes += bytes_;
        *p = static_cast<P*>(q);
        return hipSuccess;
    }

    StepArgs args() const
    {
        StepArgs A;
        std::memset(&A, 0, sizeof(A));
        A.pos = pos; A.vel = vel; A.acc = acc; A.ep = ep; A.ev = ev; A.ea = ea;
        A.force = force; A.moment = moment; A.mass = mass; A.inertia = inertia; A.tmax = tmax; A.rres = rres;
        A.scene_of = scene_of; A.scenes = d_scenes; A.walls = d_walls;
        A.cyls = d_cyls; A.cy_count = cy_count; A.cy_id = cy_id; A.cy_tsd = cy_tsd;
        A.keys = keys; A.svals = svals; A.slot_tmp = slot_tmp;
        A.bucket_count = bucket_count; A.bucket_start = bucket_start; A.scan_part = scan_part;
        A.nbr_count = nbr_count; A.nbr = nbr; A.list_cap = list_cap; A.work = work;
        A.seg_count = seg_count; A.seg_cap = seg_cap; A.work2 = work2; A.seg_count2 = seg_count2;
        A.grid_bar = grid_bar; A.rebuild_blocks = rebuild_blocks; A.rebuild_wait_limit = cfg.rebuild_wait_limit;
        A.pass_mask = pass_mask;
        A.ext_id = d_ext; A.int_id = d_int; A.jflat = d_jflat;
        A.sposf = sposf;
        A.rebuild_cur = rebuild + rbp; A.rebuild_next = rebuild + (rbp ^ 1); A.n_rebuild = n_rebuild;
        A.lctl = lctl;
        A.gap = gap; A.disp = disp;
        A.dmax_cur = dmax + (size_t)rbp * kQueueSegs * kSegStride;
        A.dmax_next = dmax + (size_t)(rbp ^ 1) * kQueueSegs * kSegStride;
        A.vel_out = vel; A.ev_out = ev;
        A.pos_next = pos_buf[spp ^ 1]; A.ep_next = ep_buf[spp ^ 1];
        A.drift = 1;
        A.pp_count_in = pp_count[cur]; A.pp_id_in = pp_id[cur]; A.pp_tsd_in = pp_tsd[cur];
        A.pp_count_out = pp_count[cur ^ 1]; A.pp_id_out = pp_id[cur ^ 1]; A.pp_tsd_out = pp_tsd[cur ^ 1];
        A.pl_count = pl_count; A.pl_id = pl_id; A.pl_tsd = pl_tsd;
        A.wrench = wrench_pp; A.err = err; A.counters = counters;
        A.meshes = d_meshes; A.tri_glob = d_tri; A.prox_glob = d_prox;
        A.mcell_start = d_mcell_start; A.mcell_tris = d_mcell_tris;
        A.tri_count = tri_count; A.tri_list = tri_list; A.msum = msum;
        A.ms_count = ms_count; A.ms_id = ms_id; A.ms_tsd = ms_tsd;
        A.mesh_work = mesh_work; A.mesh_count = mesh_count;
        A.ppsum = ppsum;
        A.has_mesh = meshes.empty() ? 0 : 1;
        A.np = np; A.nsorted = np; A.cap = cap; A.nwslot = nwslot; A.nbucket = nbucket;
        A.stale_ft = cfg.quirks.stale_ft; A.flat_inertia = cfg.quirks.flat_inertia;
        A.keep_velocity = cfg.quirks.velocity_not_updated; A.tmax_persist = tmax_persist;
        A.zero_force = 1;
        A.full_blocks = cfg.full_blocks ? std::max(cfg.full_blocks, kQueueSegs) : 1024u;
        A.resident_blocks = resident_blocks;
        A.merge_below = cfg.merge_pairs_below ? cfg.merge_pairs_below : kMergePairsBelow;
        A.has_cyl = cyls.empty() ? 0 : 1;
        return A;
    }

    // First half of a step: position update + key, bucket sort + sorted copies, pair
    // search, contacts. ev (optional) gets: start, grid done, pairs done, contacts done.
    hipError_t collide_half(bool fuse, hipEvent_t* ev)
    {
        hipError_t e;
        if (ev && (e = hipEventRecord(ev[0], stream)) != hipSuccess) return e;
        rbp ^= 1;
        if (!drift_valid && (e = run_drift()) != hipSuccess) return e;
        set_pos_parity(spp ^ 1);   // the drifted positions become current
        StepArgs A = args();
        A.drift = fuse ? 1 : 0;    // split mode: k_finish computes the next drift
        if (fuse) {
            A.vel_out = vel_buf[sv ^ 1];
            A.ev_out = ev_buf[sv ^ 1];
        }
        if ((e = launch_grid(A, multi, stream)) != hipSuccess) return e;
        if (ev && (e = hipEventRecord(ev[1], stream)) != hipSuccess) return e;
        if ((e = launch_touch(A, multi, stream)) != hipSuccess) return e;
        if (ev && (e = hipEventRecord(ev[2], stream)) != hipSuccess) return e;
        // the full pass (spheres k_touch queued) on the second stream, concurrently with the
        // lite pass (all others) on the first; the first stream waits for both: the triangle
        // pass, the pair pass (on a small full pass it returns at once: k_collide walks the
        // pairs; on the second stream its launch overlaps the lite pass instead of lengthening
        // the step), the contact kernel
        const bool gpu_law = cfg.quirks.gpu_force_law != 0;
        if ((e = hipEventRecord(fork_ev, stream)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(stream2, fork_ev, 0)) != hipSuccess) return e;
        if (!meshes.empty() && (e = collide_fn(A, multi, fuse, model, gpu_law, 2, stream2)) != hipSuccess) return e;
        if (ev && (e = hipEventRecord(ev[9], stream2)) != hipSuccess) return e;
        if ((e = collide_fn(A, multi, fuse, model, gpu_law, 3, stream2)) != hipSuccess) return e;
        if (ev && (e = hipEventRecord(ev[7], stream2)) != hipSuccess) return e;
        if ((e = collide_fn(A, multi, fuse, model, gpu_law, 1, stream2)) != hipSuccess) return e;
        if (ev && (e = hipEventRecord(ev[8], stream2)) != hipSuccess) return e;
        if ((e = hipEventRecord(join_ev, stream2)) != hipSuccess) return e;
        if (ev && fuse) {
            // timed steps: k_lite_move bracketed by its own events
            if ((e = hipEventRecord(ev[5], stream)) != hipSuccess) return e;
            launch_lite_fused(A, multi, stream, 1);
            if ((e = hipEventRecord(ev[6], stream)) != hipSuccess) return e;
            launch_lite_fused(A, multi, stream, 2);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        } else {
            if (ev && ((e = hipEventRecord(ev[5], stream)) != hipSuccess || (e = hipEventRecord(ev[6], stream)) != hipSuccess))
                return e;
            if ((e = collide_fn(A, multi, fuse, model, gpu_law, 0, stream)) != hipSuccess) return e;
        }
        if ((e = hipStreamWaitEvent(stream, join_ev, 0)) != hipSuccess) return e;
        if (ev && (e = hipEventRecord(ev[3], stream)) != hipSuccess) return e;
        if (nwslot) {
            hipLaunchKernelGGL(k_wrench_partial, dim3(kWrenchBlocks, nwslot), dim3(256), 0, stream, wrench_pp,
                               nbr_count, np, d_wslots, wrench_partial);
            hipLaunchKernelGGL(k_wrench_final, dim3(nwslot), dim3(kWrenchBlocks), 0, stream, wrench_partial,
                               wrench_out);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        cur ^= 1;
        if (fuse) set_state_parity(sv ^ 1);
        drift_valid = fuse;
        return hipSuccess;
    }

    void set_state_parity(int p)
    {
        sv = p;
        vel = vel_buf[p];
        ev = ev_buf[p];
    }

    // Drift of the current state into the next buffers (inside a step, after its rbp flip),
    // and a list rebuild for this step: the lists' reference positions are those of an
    // older state.
    hipError_t run_drift()
    {
        StepArgs A = args();
        A.rebuild_next = nullptr;
        A.dmax_next = nullptr;
        hipError_t e = launch_drift(A, multi, stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, stream, rebuild + rbp, 1u, (size_t)1);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        drift_valid = true;
        return hipSuccess;
    }

    // Between steps: a wall or mesh body moved by at most d (its largest point displacement);
    // the next step rebuilds the lists once the summed motion since the last build reaches
    // half the smallest skin.
    hipError_t body_moved(double d)
    {
        hipLaunchKernelGGL(k_body_moved, dim3(1), dim3(1), 0, stream, lctl, d, rebuild + (rbp ^ 1));
        return hipGetLastError();
    }

    void set_pos_parity(int p)
    {
        spp = p;
        pos = pos_buf[p];
        ep = ep_buf[p];
    }

    // Rolling resistance + velocity update (+ the next drift) of a split step.
    hipError_t finish_half(bool zero_force)
    {
        StepArgs A = args();
        A.zero_force = zero_force ? 1 : 0;
        const hipError_t e = launch_finish(A, multi, stream);
        if (e == hipSuccess) drift_valid = true;
        return e;
    }

    // One full step; ev (optional, kEventsPerStep) brackets grid, pair search, contacts and
    // the velocity kernel.
    hipError_t full_step(hipEvent_t* ev)
    {
        hipError_t e;
        if (cfg.fuse_velocity) {
            if ((e = collide_half(true, ev)) != hipSuccess) return e;
            if (ev) return hipEventRecord(ev[4], stream);
            return hipSuccess;
        }
        if ((e = collide_half(false, ev)) != hipSuccess) return e;
        // the next contact pass overwrites force and moment: no need to clear them
        if ((e = finish_half(false)) != hipSuccess) return e;
        if (ev) return hipEventRecord(ev[4], stream);
        return hipSuccess;
    }
};

constexpr unsigned kEventsPerStep = 10;

static int hip_status(hipError_t e) { return e == hipSuccess ? XDEM_OK : XDEM_E_HIP; }

#define CTX_CHECK(expr)                                  \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return XDEM_E_HIP;         \
    } while (0)

extern "C" {

int xdem_ctx_create(const xdem_ctx_config* cfg, xdem_ctx** out)
{
    if (!out) return XDEM_E_INVALID;
    *out = nullptr;
    xdem_ctx* c = new (std::nothrow) xdem_ctx();
    if (!c) return XDEM_E_INVALID;
    if (cfg) c->cfg = *cfg;
    else {
        std::memset(&c->cfg, 0, sizeof(c->cfg));
        c->cfg.quirks.stale_ft = 1;
        c->cfg.quirks.flat_inertia = 1;
    }
    if (c->cfg.precision != XDEM_PRECISION_FP64 && c->cfg.precision != XDEM_PRECISION_MIXED) { delete c; return XDEM_E_INVALID; }
    c->cap = c->cfg.p2p_capacity ? c->cfg.p2p_capacity : XDEM_MAX_P2P_COUNT;
    if (c->cap > 64) { delete c; return XDEM_E_INVALID; }
    c->auto_steps = c->cfg.step_graph > 0 ? (unsigned)(c->cfg.step_graph + 1) & ~1u : 0u;
    if (!c->cfg.grid_dim_max) c->cfg.grid_dim_max = 512;
    if (hipSetDevice(c->cfg.device) != hipSuccess) { delete c; return XDEM_E_HIP; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming) != hipSuccess) {
        xdem_ctx_destroy(c);
        return XDEM_E_HIP;
    }
    *out = c;
    return XDEM_OK;
}

int xdem_ctx_destroy(xdem_ctx* c)
{
    if (!c) return XDEM_OK;
    // teardown: a failing call leaves nothing to undo, so the codes are not checked
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);   // the second stream joins into it
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->agexec) (void)hipGraphExecDestroy(c->agexec);
    if (c->agraph) (void)hipGraphDestroy(c->agraph);
    for (int q = 0; q < 2; q++) {
        if (c->sgexec[q]) (void)hipGraphExecDestroy(c->sgexec[q]);
        if (c->sgraph[q]) (void)hipGraphDestroy(c->sgraph[q]);
    }
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->h_wrench) (void)hipHostFree(c->h_wrench);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->join_ev) (void)hipEventDestroy(c->join_ev);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return XDEM_OK;
}

int xdem_ctx_add_scene(xdem_ctx* c, const xdem_scene* s, unsigned int* idx)
{
    if (!c || !s) return XDEM_E_INVALID;
    if (c->committed) return XDEM_E_NOT_READY;
    if (s->np == 0 || !s->pos || !s->mass || !s->inertia) return XDEM_E_INVALID;
    if (s->nplane > (unsigned)kMaxWallsPerScene) return XDEM_E_INVALID;
    if (s->nplane && (!s->planes || !s->plane_dbi || !s->plane_cp)) return XDEM_E_INVALID;
    if (!(s->cell_size > 0.0) || !(s->dt > 0.0)) return XDEM_E_INVALID;
    if (c->model >= 0 && c->model != s->contact_model) return XDEM_E_INVALID;  // one force model per run
    if ((unsigned long long)c->np + s->np >= 0xfffffff0ull) return XDEM_E_INVALID;
    for (unsigned k = 0; k < s->nplane; k++)
        if (s->planes[k].id != k) return XDEM_E_INVALID;  // wall ids follow the wall order
    c->model = s->contact_model;
    SceneHost h;
    h.s = *s;
    h.pos.assign(s->pos, s->pos + 4 * (size_t)s->np);
    h.vel.assign(3 * (size_t)s->np, 0.0);
    if (s->vel) h.vel.assign(s->vel, s->vel + 3 * (size_t)s->np);
    h.mass.assign(s->mass, s->mass + s->np);
    h.inertia.assign(s->inertia, s->inertia + 3 * (size_t)s->np);
    h.planes.assign(s->planes, s->planes + s->nplane);
    h.plane_cp.assign(s->plane_cp, s->plane_cp + s->nplane);
    h.dbi.assign(s->plane_dbi, s->plane_dbi + 15 * (size_t)s->nplane);
    // local corners = A(e)^T (corner - body position) at the pose given with the scene
    // (xPlaneObject::define, source/xdynamics_object/xPlaneObject.cpp:45-48)
    h.wall_local.assign(12 * (size_t)s->nplane, 0.0);
    for (unsigned k = 0; k < s->nplane; k++) {
        const double* d = s->plane_dbi + 15 * k;
        const xdem_double3 cs[4] = {s->planes[k].xw, s->planes[k].w2, s->planes[k].w3, s->planes[k].w4};
        for (int q = 0; q < 4; q++) {
            const double rel[3] = {cs[q].x - d[XDEM_PX], cs[q].y - d[XDEM_PY], cs[q].z - d[XDEM_PZ]};
            euler_rotate(d + XDEM_E0, rel, true, &h.wall_local[12 * k + 3 * q]);
        }
    }
    h.p_begin = c->np;
    h.wall_begin = (uint32_t)c->walls.size();
    c->np += s->np;
    double rmax = 0.0, rmin = 1e300;
    for (unsigned i = 0; i < s->np; i++) {
        rmax = std::max(rmax, h.pos[4 * i + 3]);
        rmin = std::min(rmin, h.pos[4 * i + 3]);
    }
    if (s->np == 0) rmin = 0.0;
    // neighbour-list skin and bucket size: buckets hold every pair within 2 r_max + reach +
    // skin in neighbouring buckets (the reference's cell size stays the one of the visiting
    // order)
    const ContactConst cc_pp = make_contact_const(s->cp_pp, 0.5 * rmax);
    // default: adaptive, up to skin_cap (0.4) of the smallest radius, starting at a sixteenth of it
    const double cap = c->cfg.skin_cap > 0.0 ? c->cfg.skin_cap : 0.4;
    const double skin = c->cfg.skin > 0.0 ? c->cfg.skin : (c->cfg.skin == 0.0 ? cap * rmin : 0.0);
    const double bucket_cs = std::max(s->cell_size, 2.0 * rmax + cc_pp.coh_reach + skin) * (1.0 + 1e-9);
    // bucket grid: the reference's origin and cell size, power-of-two dims covering the
    // scene (wrapping beyond them, as the reference's 128^3 grid does)
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (unsigned i = 0; i < s->np; i++)
        for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], h.pos[4 * i + a]); hi[a] = std::max(hi[a], h.pos[4 * i + a]); }
    for (unsigned k = 0; k < s->nplane; k++) {
        const xdem_double3 cs[4] = {s->planes[k].xw, s->planes[k].w2, s->planes[k].w3, s->planes[k].w4};
        for (const auto& q : cs) {
            const double v[3] = {q.x, q.y, q.z};
            for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], v[a]); hi[a] = std::max(hi[a], v[a]); }
        }
    }
    const double wo[3] = {s->world_origin[0], s->world_origin[1], s->world_origin[2]};
    uint32_t dims[3];
    for (int a = 0; a < 3; a++) {
        const double span = std::floor((hi[a] - wo[a]) / bucket_cs) - std::floor((lo[a] - wo[a]) / bucket_cs) + 3.0;
        uint32_t d = next_pow2((uint32_t)std::min(span, 1048576.0));
        d = std::max(4u, std::min(d, c->cfg.grid_dim_max));
        dims[a] = d;
    }
    // Bucket budget: halve the largest axis (wrapping it, as the reference's 128^3 grid
    // wraps) until the scene's grid has at most next_pow2(buckets_per_sphere * np) buckets.
    // Every axis keeps at least 4 buckets, so the 27 neighbours of a cell stay 27 distinct
    // buckets and the visiting order of real contacts does not change.
    {
        const unsigned long long per = c->cfg.buckets_per_sphere ? c->cfg.buckets_per_sphere : 4u;
        unsigned long long budget = 64;
        while (budget < per * (unsigned long long)std::max(1u, s->np)) budget <<= 1;
        while ((unsigned long long)dims[0] * dims[1] * dims[2] > budget) {
            int a = 0;
            if (dims[1] > dims[a]) a = 1;
            if (dims[2] > dims[a]) a = 2;
            if (dims[a] <= 4) break;
            dims[a] >>= 1;
        }
    }
    SceneDev& D = h.dev;
    std::memset(&D, 0, sizeof(D));
    D.dt = s->dt;
    D.half2dt = 0.5 * s->dt * s->dt;
    D.gx = s->gravity[0]; D.gy = s->gravity[1]; D.gz = s->gravity[2];
    D.grid.wo_x = wo[0]; D.grid.wo_y = wo[1]; D.grid.wo_z = wo[2];
    D.grid.cs = bucket_cs;
    D.ref_cs = s->cell_size;
    D.skin = skin;
    D.grid.mask_x = dims[0] - 1; D.grid.mask_y = dims[1] - 1; D.grid.mask_z = dims[2] - 1;
    D.grid.dim_x = dims[0];
    D.grid.dim_xy = dims[0] * dims[1];
    D.grid.base = c->nbucket;
    const unsigned long long nb = (unsigned long long)dims[0] * dims[1] * dims[2];
    if ((unsigned long long)c->nbucket + nb >= 0xffffffffull) return XDEM_E_INVALID;
    c->nbucket += (uint32_t)nb;
    D.cc = cc_pp;
    D.p_begin = h.p_begin;
    D.p_end = h.p_begin + s->np;
    D.wall_begin = h.wall_begin;
    D.nwall = s->nplane;
    D.has_pp = 1;
    D.mesh = -1;
    h.rmax = rmax;
    if (s->cp_pp.rfric == 0.0) c->tmax_persist = 1;
    for (unsigned k = 0; k < s->nplane; k++) {
        WallDev W;
        std::memset(&W, 0, sizeof(W));
        W.info = s->planes[k];
        std::memcpy(W.dbi, s->plane_dbi + 15 * k, sizeof(double) * 15);
        W.cc = make_contact_const(s->plane_cp[k], rmax);
        if (s->plane_cp[k].rfric == 0.0) c->tmax_persist = 1;
        W.wslot = -1;
        if (W.dbi[XDEM_MM] > 0.0) {  // a wall with mass is a dynamic body: reduce its wrench
            W.wslot = (int32_t)c->nwslot++;
            c->wslots.push_back({D.p_begin, D.p_end, kTouch, 0u});
        }
        if (W.info.id == 0) D.wall_id0 = 1;
        if (W.wslot >= 0) D.wall_dynamic = 1;
        c->wall_slot.push_back(W.wslot);
        c->walls.push_back(W);
    }
    if (idx) *idx = (unsigned)c->scenes.size();
    c->scenes.push_back(std::move(h));
    return XDEM_OK;
}

int xdem_ctx_commit(xdem_ctx* c)
{
    if (!c) return XDEM_E_INVALID;
    if (c->committed) return XDEM_E_NOT_READY;
    if (c->scenes.empty()) return XDEM_E_INVALID;
    CTX_CHECK(hipSetDevice(c->cfg.device));
    if (c->np >= (1u << kIdBits)) return XDEM_E_INVALID;   // ids packed below the visiting-order code
    c->multi = c->scenes.size() > 1;
    c->collide_fn = c->cfg.precision == XDEM_PRECISION_MIXED ? collide_launch_mixed() : collide_launch_f64();
    const size_t n = c->np;
    c->key_bits = 1;
    while (c->key_bits < 32 && (1ull << c->key_bits) < (unsigned long long)c->nbucket) c->key_bits++;
    CTX_CHECK(c->alloc(&c->pos_buf[0], sizeof(double) * 4 * n));
    CTX_CHECK(c->alloc(&c->pos_buf[1], sizeof(double) * 4 * n));
    CTX_CHECK(c->alloc(&c->vel_buf[0], sizeof(double) * 3 * n));
    if (c->cfg.fuse_velocity) CTX_CHECK(c->alloc(&c->vel_buf[1], sizeof(double) * 3 * n));
    c->vel = c->vel_buf[0];
    CTX_CHECK(c->alloc(&c->acc, sizeof(double) * 3 * n));
    CTX_CHECK(c->alloc(&c->ep_buf[0], sizeof(double) * 4 * n));
    CTX_CHECK(c->alloc(&c->ep_buf[1], sizeof(double) * 4 * n));
    c->set_pos_parity(0);
    c->drift_valid = false;
    CTX_CHECK(c->alloc(&c->ev_buf[0], sizeof(double) * 4 * n));
    if (c->cfg.fuse_velocity) CTX_CHECK(c->alloc(&c->ev_buf[1], sizeof(double) * 4 * n));
    c->ev = c->ev_buf[0];
    CTX_CHECK(c->alloc(&c->ea, sizeof(double) * 4 * n));
    CTX_CHECK(c->alloc(&c->force, sizeof(double) * 3 * n));
    CTX_CHECK(c->alloc(&c->moment, sizeof(double) * 3 * n));
    CTX_CHECK(c->alloc(&c->mass, sizeof(double) * n));
    CTX_CHECK(c->alloc(&c->inertia, sizeof(double) * 3 * n));
    CTX_CHECK(c->alloc(&c->tmax, sizeof(double) * 3 * n));
    CTX_CHECK(c->alloc(&c->rres, sizeof(double) * n));
    if (c->multi) CTX_CHECK(c->alloc(&c->scene_of, sizeof(uint32_t) * n));
    CTX_CHECK(c->alloc(&c->d_scenes, sizeof(SceneDev) * c->scenes.size()));
    CTX_CHECK(c->alloc(&c->d_walls, sizeof(WallDev) * std::max<size_t>(1, c->walls.size())));
    CTX_CHECK(c->alloc(&c->keys, sizeof(uint32_t) * n));
    CTX_CHECK(c->alloc(&c->svals, sizeof(uint32_t) * n));
    CTX_CHECK(c->alloc(&c->slot_tmp, sizeof(uint32_t) * n));
    c->list_cap = std::min(64u, std::max(48u, 2u * c->cap));   // k_touch's pass mask has 64 bits
    CTX_CHECK(c->alloc(&c->nbr_count, sizeof(uint32_t) * n));
    c->seg_cap = queue_segment_cap((uint32_t)n);
    CTX_CHECK(c->alloc(&c->work, sizeof(uint32_t) * kQueueSegs * c->seg_cap));
    CTX_CHECK(c->alloc(&c->seg_count, sizeof(uint32_t) * kQueueSegs * kSegStride));
    CTX_CHECK(c->alloc(&c->work2, sizeof(uint32_t) * kQueueSegs * c->seg_cap));
    CTX_CHECK(c->alloc(&c->pass_mask, sizeof(unsigned long long) * n));
    CTX_CHECK(c->alloc(&c->d_ext, sizeof(uint32_t) * n));
    CTX_CHECK(c->alloc(&c->d_int, sizeof(uint32_t) * n));
    CTX_CHECK(c->alloc(&c->d_jflat, sizeof(double) * n));
    CTX_CHECK(c->alloc(&c->seg_count2, sizeof(uint32_t) * kQueueSegs * kSegStride));
    CTX_CHECK(c->alloc(&c->grid_bar, sizeof(uint32_t) * kSegStride * kBarSlots));
    c->rebuild_blocks = rebuild_grid_size(c->multi);
    {
        int cus = 0;
        CTX_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->cfg.device));
        c->resident_blocks = resident_grid(cus);
    }
    if (c->rebuild_blocks == 0) return XDEM_E_HIP;
    CTX_CHECK(c->alloc(&c->nbr, sizeof(uint32_t) * n * c->list_cap));
    const size_t nbt = bucket_table_size(c->nbucket);
    CTX_CHECK(c->alloc(&c->bucket_count, sizeof(uint32_t) * nbt));
    CTX_CHECK(c->alloc(&c->bucket_start, sizeof(uint32_t) * (nbt + 1)));
    CTX_CHECK(c->alloc(&c->scan_part, sizeof(uint32_t) * (nbt / 4096 + 1)));
    CTX_CHECK(c->alloc(&c->sposf, sizeof(float4) * n));
    CTX_CHECK(c->alloc(&c->rebuild, sizeof(uint32_t) * 2));
    CTX_CHECK(c->alloc(&c->lctl, sizeof(ListCtl)));
    CTX_CHECK(c->alloc(&c->dmax, sizeof(uint32_t) * 2 * kQueueSegs * kSegStride));
    CTX_CHECK(c->alloc(&c->gap, sizeof(float) * n));
    CTX_CHECK(c->alloc(&c->disp, sizeof(float) * n));
    CTX_CHECK(c->alloc(&c->n_rebuild, sizeof(unsigned long long)));
    for (int b = 0; b < 2; b++) {
        CTX_CHECK(c->alloc(&c->pp_count[b], sizeof(uint32_t) * n));
        CTX_CHECK(c->alloc(&c->pp_id[b], sizeof(uint32_t) * n * c->cap));
        CTX_CHECK(c->alloc(&c->pp_tsd[b], sizeof(double2) * n * c->cap));
    }
    CTX_CHECK(c->alloc(&c->ppsum, sizeof(double) * 10 * n));
    CTX_CHECK(c->alloc(&c->pl_count, sizeof(uint32_t) * n));
    CTX_CHECK(c->alloc(&c->pl_id, sizeof(uint32_t) * n * XDEM_MAX_P2PL_COUNT));
    CTX_CHECK(c->alloc(&c->pl_tsd, sizeof(double2) * n * XDEM_MAX_P2PL_COUNT));
    std::vector<CylDev> cdev;
    if (!c->cyls.empty()) {
        // cylinder bodies grouped by scene, each scene's in the order added; their slots
        // (XDEM_MAX_P2CY_COUNT per sphere, updated in place)
        c->cyl_dev.assign(c->cyls.size(), 0u);
        for (size_t si = 0; si < c->scenes.size(); si++) {
            SceneDev& D = c->scenes[si].dev;
            D.cyl_begin = (uint32_t)cdev.size();
            D.ncyl = 0;
            D.cyl_id0 = 0;
            for (size_t k = 0; k < c->cyls.size(); k++) {
                if (c->cyls[k].scene != si) continue;
                c->cyl_dev[k] = (uint32_t)cdev.size();
                cdev.push_back(c->cyls[k].dev);
                D.ncyl++;
                if (c->cyls[k].dev.info.id == 0) D.cyl_id0 = 1;
            }
        }
        CTX_CHECK(c->alloc(&c->d_cyls, sizeof(CylDev) * cdev.size()));
        CTX_CHECK(c->alloc(&c->cy_count, sizeof(uint32_t) * n));
        CTX_CHECK(c->alloc(&c->cy_id, sizeof(uint32_t) * n * XDEM_MAX_P2CY_COUNT));
        CTX_CHECK(c->alloc(&c->cy_tsd, sizeof(double2) * n * XDEM_MAX_P2CY_COUNT));
    }
    if (c->nwslot) {
        CTX_CHECK(c->alloc(&c->wrench_pp, sizeof(double) * 6 * n * c->nwslot));
        CTX_CHECK(c->alloc(&c->wrench_out, sizeof(double) * 6 * c->nwslot));
        CTX_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c->h_wrench), sizeof(double) * 6, hipHostMallocDefault));
        CTX_CHECK(c->alloc(&c->wrench_partial, sizeof(double) * 6 * kWrenchBlocks * c->nwslot));
        CTX_CHECK(c->alloc(&c->d_wslots, sizeof(WrenchSlot) * c->nwslot));
    }
    std::vector<MeshDev> mdev;
    std::vector<double> all_vlist, all_plocal, all_prad;
    std::vector<uint32_t> all_cell_start, all_cell_tris;
    if (!c->meshes.empty()) {
        // mesh bodies: body-frame data, the grids' cell tables concatenated (each mesh's cell
        // starts offset by its first entry), per-sphere triangle lists, sums and history
        for (MeshHost& m : c->meshes) {
            m.dev.cell_base = (uint32_t)all_cell_start.size();
            const uint32_t off = (uint32_t)all_cell_tris.size();
            for (size_t k = 0; k + 1 < m.cell_start.size(); k++) all_cell_start.push_back(off + m.cell_start[k]);
            all_cell_tris.insert(all_cell_tris.end(), m.cell_tris.begin(), m.cell_tris.end());
            all_vlist.insert(all_vlist.end(), m.vlist.begin(), m.vlist.end());
            all_plocal.insert(all_plocal.end(), m.plocal.begin(), m.plocal.end());
            all_prad.insert(all_prad.end(), m.prad.begin(), m.prad.end());
            mdev.push_back(m.dev);
        }
        all_cell_start.push_back((uint32_t)all_cell_tris.size());
        const size_t T = c->ntri;
        CTX_CHECK(c->alloc(&c->d_meshes, sizeof(MeshDev) * mdev.size()));
        CTX_CHECK(c->alloc(&c->d_vlist, sizeof(double) * 9 * T));
        CTX_CHECK(c->alloc(&c->d_plocal, sizeof(double) * 3 * T));
        CTX_CHECK(c->alloc(&c->d_prad, sizeof(double) * T));
        CTX_CHECK(c->alloc(&c->d_prox, sizeof(double) * 4 * T));
        CTX_CHECK(c->alloc(&c->d_tri, sizeof(xdem_triangle_info) * T));
        CTX_CHECK(c->alloc(&c->d_mcell_start, sizeof(uint32_t) * all_cell_start.size()));
        CTX_CHECK(c->alloc(&c->d_mcell_tris, sizeof(uint32_t) * std::max<size_t>(1, all_cell_tris.size())));
        CTX_CHECK(c->alloc(&c->tri_count, sizeof(uint32_t) * n));
        CTX_CHECK(c->alloc(&c->tri_list, sizeof(uint32_t) * n * kTriCap));
        CTX_CHECK(c->alloc(&c->msum, sizeof(double) * 10 * n));
        CTX_CHECK(c->alloc(&c->ms_count, sizeof(uint32_t) * n));
        CTX_CHECK(c->alloc(&c->ms_id, sizeof(uint32_t) * n * XDEM_MAX_P2MS_COUNT));
        CTX_CHECK(c->alloc(&c->ms_tsd, sizeof(double2) * n * XDEM_MAX_P2MS_COUNT));
        CTX_CHECK(c->alloc(&c->mesh_work, sizeof(uint32_t) * n));
        CTX_CHECK(c->alloc(&c->mesh_count, sizeof(uint32_t)));
    }
    CTX_CHECK(c->alloc(&c->err, sizeof(uint32_t)));
    CTX_CHECK(c->alloc(&c->counters, sizeof(unsigned long long) * 4));

    // host staging of the initial state (xDiscreteElementMethodSimulation.cpp:60-83)
    std::vector<double> pos(4 * n), vel(3 * n), acc(3 * n, 0.0), ep(4 * n, 0.0), mass(n), inertia(3 * n);
    std::vector<uint32_t> scene_of(c->multi ? n : 0);
    std::vector<double> jflat(n);
    std::vector<SceneDev> sd;
    c->ext_of.assign(n, 0u);
    c->int_of.assign(n, 0u);
    for (size_t si = 0; si < c->scenes.size(); si++) {
        const SceneHost& h = c->scenes[si];
        const size_t b = h.p_begin, m = h.s.np;
        // stored order: the caller's, or a stable sort on the bucket key of the initial
        // position (the order of the grid's sorted slots at the first build)
        std::vector<uint32_t> ord(m);
        std::iota(ord.begin(), ord.end(), 0u);
        if (!c->cfg.keep_order) {
            const Grid& g = h.dev.grid;
            std::vector<uint32_t> key(m);
            for (size_t e = 0; e < m; e++) {
                const int gx = (int)std::floor((h.pos[4 * e] - g.wo_x) / g.cs);
                const int gy = (int)std::floor((h.pos[4 * e + 1] - g.wo_y) / g.cs);
                const int gz = (int)std::floor((h.pos[4 * e + 2] - g.wo_z) / g.cs);
                key[e] = ((uint32_t)gz & g.mask_z) * g.dim_xy + ((uint32_t)gy & g.mask_y) * g.dim_x +
                         ((uint32_t)gx & g.mask_x);
            }
            std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return key[x] < key[y]; });
        }
        for (size_t j = 0; j < m; j++) {
            const size_t e = ord[j], i = b + j;
            c->ext_of[i] = (uint32_t)(b + e);
            c->int_of[b + e] = (uint32_t)i;
            std::copy(h.pos.begin() + 4 * e, h.pos.begin() + 4 * e + 4, pos.begin() + 4 * i);
            std::copy(h.vel.begin() + 3 * e, h.vel.begin() + 3 * e + 3, vel.begin() + 3 * i);
            mass[i] = h.mass[e];
            std::copy(h.inertia.begin() + 3 * e, h.inertia.begin() + 3 * e + 3, inertia.begin() + 3 * i);
            // quirk Q5 reads the scene's inertia array as flat doubles at the caller's index
            jflat[i] = h.inertia[e];
            acc[3 * i + 1] = h.s.gravity[1];  // acc = (0, g.y, 0) as the reference initialises it
            ep[4 * i] = 1.0;
            if (c->multi) scene_of[i] = (uint32_t)si;
        }
        sd.push_back(h.dev);
    }
    hipStream_t s = c->stream;
    CTX_CHECK(hipMemcpyAsync(c->pos, pos.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice, s));
    CTX_CHECK(hipMemcpyAsync(c->vel, vel.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    CTX_CHECK(hipMemcpyAsync(c->acc, acc.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    CTX_CHECK(hipMemcpyAsync(c->ep, ep.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice, s));
    CTX_CHECK(hipMemcpyAsync(c->mass, mass.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    CTX_CHECK(hipMemcpyAsync(c->inertia, inertia.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    CTX_CHECK(hipMemcpyAsync(c->d_ext, c->ext_of.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    CTX_CHECK(hipMemcpyAsync(c->d_int, c->int_of.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    CTX_CHECK(hipMemcpyAsync(c->d_jflat, jflat.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (c->multi) CTX_CHECK(hipMemcpyAsync(c->scene_of, scene_of.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    CTX_CHECK(hipMemcpyAsync(c->d_scenes, sd.data(), sizeof(SceneDev) * sd.size(), hipMemcpyHostToDevice, s));
    if (!c->walls.empty())
        CTX_CHECK(hipMemcpyAsync(c->d_walls, c->walls.data(), sizeof(WallDev) * c->walls.size(), hipMemcpyHostToDevice, s));
    if (c->nwslot)
        CTX_CHECK(hipMemcpyAsync(c->d_wslots, c->wslots.data(), sizeof(WrenchSlot) * c->nwslot, hipMemcpyHostToDevice,
                                 s));
    if (!cdev.empty()) {
        CTX_CHECK(hipMemcpyAsync(c->d_cyls, cdev.data(), sizeof(CylDev) * cdev.size(), hipMemcpyHostToDevice, s));
        CTX_CHECK(hipMemsetAsync(c->cy_count, 0, sizeof(uint32_t) * n, s));
        CTX_CHECK(hipMemsetAsync(c->cy_id, 0, sizeof(uint32_t) * n * XDEM_MAX_P2CY_COUNT, s));
        CTX_CHECK(hipMemsetAsync(c->cy_tsd, 0, sizeof(double2) * n * XDEM_MAX_P2CY_COUNT, s));
    }
    if (!c->meshes.empty()) {
        const size_t T = c->ntri;
        CTX_CHECK(hipMemcpyAsync(c->d_meshes, mdev.data(), sizeof(MeshDev) * mdev.size(), hipMemcpyHostToDevice, s));
        CTX_CHECK(hipMemcpyAsync(c->d_vlist, all_vlist.data(), sizeof(double) * 9 * T, hipMemcpyHostToDevice, s));
        CTX_CHECK(hipMemcpyAsync(c->d_plocal, all_plocal.data(), sizeof(double) * 3 * T, hipMemcpyHostToDevice, s));
        CTX_CHECK(hipMemcpyAsync(c->d_prad, all_prad.data(), sizeof(double) * T, hipMemcpyHostToDevice, s));
        CTX_CHECK(hipMemcpyAsync(c->d_mcell_start, all_cell_start.data(), sizeof(uint32_t) * all_cell_start.size(),
                                 hipMemcpyHostToDevice, s));
        if (!all_cell_tris.empty())
            CTX_CHECK(hipMemcpyAsync(c->d_mcell_tris, all_cell_tris.data(), sizeof(uint32_t) * all_cell_tris.size(),
                                     hipMemcpyHostToDevice, s));
        CTX_CHECK(hipMemsetAsync(c->tri_count, 0, sizeof(uint32_t) * n, s));
        CTX_CHECK(hipMemsetAsync(c->ms_count, 0, sizeof(uint32_t) * n, s));
        CTX_CHECK(hipMemsetAsync(c->ms_id, 0, sizeof(uint32_t) * n * XDEM_MAX_P2MS_COUNT, s));
        CTX_CHECK(hipMemsetAsync(c->ms_tsd, 0, sizeof(double2) * n * XDEM_MAX_P2MS_COUNT, s));
        CTX_CHECK(hipMemsetAsync(c->mesh_count, 0, sizeof(uint32_t), s));
        for (uint32_t m = 0; m < c->meshes.size(); m++) {
            hipLaunchKernelGGL(k_mesh_pose, dim3((c->meshes[m].dev.ntri + 255) / 256), dim3(256), 0, s, c->d_meshes, m,
                               c->d_vlist, c->d_plocal, c->d_prad, c->d_tri, c->d_prox);
            CTX_CHECK(hipGetLastError());
        }
    }
    CTX_CHECK(hipMemsetAsync(c->ev, 0, sizeof(double) * 4 * n, s));
    CTX_CHECK(hipMemsetAsync(c->ea, 0, sizeof(double) * 4 * n, s));
    CTX_CHECK(hipMemsetAsync(c->force, 0, sizeof(double) * 3 * n, s));
    CTX_CHECK(hipMemsetAsync(c->moment, 0, sizeof(double) * 3 * n, s));
    CTX_CHECK(hipMemsetAsync(c->tmax, 0, sizeof(double) * 3 * n, s));
    CTX_CHECK(hipMemsetAsync(c->rres, 0, sizeof(double) * n, s));
    for (int b = 0; b < 2; b++) {
        CTX_CHECK(hipMemsetAsync(c->pp_count[b], 0, sizeof(uint32_t) * n, s));
        CTX_CHECK(hipMemsetAsync(c->pp_id[b], 0, sizeof(uint32_t) * n * c->cap, s));
        CTX_CHECK(hipMemsetAsync(c->pp_tsd[b], 0, sizeof(double2) * n * c->cap, s));
    }
    CTX_CHECK(hipMemsetAsync(c->pl_count, 0, sizeof(uint32_t) * n, s));
    CTX_CHECK(hipMemsetAsync(c->pl_id, 0, sizeof(uint32_t) * n * XDEM_MAX_P2PL_COUNT, s));
    CTX_CHECK(hipMemsetAsync(c->pl_tsd, 0, sizeof(double2) * n * XDEM_MAX_P2PL_COUNT, s));
    CTX_CHECK(hipMemsetAsync(c->err, 0, sizeof(uint32_t), s));
    CTX_CHECK(hipMemsetAsync(c->counters, 0, sizeof(unsigned long long) * 4, s));
    CTX_CHECK(hipMemsetAsync(c->n_rebuild, 0, sizeof(unsigned long long), s));
    {
        ListCtl L;
        std::memset(&L, 0, sizeof(L));
        L.adaptive = c->cfg.skin == 0.0 ? 1u : 0u;
        // the adaptive skin starts at its smallest value, a sixteenth of the cap: short runs
        // keep short lists, long ones grow the skin as builds come too often
        L.scale = L.adaptive ? 0.0625f : 1.0f;
        L.min_scale = L.adaptive ? 0.0625f : 1.0f;
        L.half_skin_min = 1e300;
        for (const SceneHost& h : c->scenes) L.half_skin_min = std::min(L.half_skin_min, 0.5 * h.dev.skin);
        CTX_CHECK(hipStreamSynchronize(s));
        CTX_CHECK(hipMemcpy(c->lctl, &L, sizeof(L), hipMemcpyHostToDevice));
    }
    CTX_CHECK(hipMemsetAsync(c->dmax, 0, sizeof(uint32_t) * 2 * kQueueSegs * kSegStride, s));
    CTX_CHECK(hipMemsetAsync(c->disp, 0, sizeof(float) * n, s));
    CTX_CHECK(hipMemsetAsync(c->gap, 0, sizeof(float) * n, s));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, c->rebuild, 1u, (size_t)2);   // the first step builds
    CTX_CHECK(hipMemsetAsync(c->seg_count, 0, sizeof(uint32_t) * kQueueSegs * kSegStride, s));
    CTX_CHECK(hipMemsetAsync(c->seg_count2, 0, sizeof(uint32_t) * kQueueSegs * kSegStride, s));
    CTX_CHECK(hipMemsetAsync(c->grid_bar, 0, sizeof(uint32_t) * kSegStride * kBarSlots, s));
    CTX_CHECK(hipMemsetAsync(c->nbr_count, 0, sizeof(uint32_t) * n, s));
    if (c->nwslot) CTX_CHECK(hipMemsetAsync(c->wrench_out, 0, sizeof(double) * 6 * c->nwslot, s));
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nbt + 255) / 256)), dim3(256), 0, s, c->bucket_count, 0u, nbt);
    CTX_CHECK(hipGetLastError());
    CTX_CHECK(hipStreamSynchronize(s));
    c->committed = true;
    return XDEM_OK;
}

// Records nsteps steps into a HIP graph without running them: the drift must be valid (it is
// computed first otherwise); the parities (history, state, position, rebuild flag) the graph's
// buffers assume are returned in par and restored afterwards. An even count leaves them as it
// found them (the graph can be replayed back to back); one step flips them all.
static int capture_steps(xdem_ctx* c, unsigned nsteps, hipGraph_t* graph, hipGraphExec_t* exec, int par[4])
{
    if (!c->drift_valid) CTX_CHECK(c->run_drift());   // the graph's steps start from a precomputed drift
    CTX_CHECK(hipStreamSynchronize(c->stream));
    par[0] = c->cur; par[1] = c->sv; par[2] = c->spp; par[3] = c->rbp;
    CTX_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (unsigned k = 0; k < nsteps && e == hipSuccess; k++) e = c->full_step(nullptr);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    c->cur = par[0];
    c->rbp = par[3];
    c->set_state_parity(par[1]);
    c->set_pos_parity(par[2]);
    c->drift_valid = true;
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);   // already failing: report the capture error
        return XDEM_E_HIP;
    }
    *graph = g;
    CTX_CHECK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    // the executable's one-time transfer to the device now, not inside the first replay
    CTX_CHECK(hipGraphUpload(*exec, c->stream));
    CTX_CHECK(hipStreamSynchronize(c->stream));
    return XDEM_OK;
}

static bool parities_match(const xdem_ctx* c, const int par[4])
{
    return c->cur == par[0] && c->sv == par[1] && c->spp == par[2] && c->rbp == par[3] && c->drift_valid;
}

static int run_steps(xdem_ctx* c, unsigned nsteps)
{
    const bool time_it = c->timing;
    const unsigned kmax = 2048u;
    if (time_it) {
        const size_t need = kEventsPerStep * (size_t)std::min(nsteps, kmax);
        while (c->events.size() < need) {
            hipEvent_t e;
            CTX_CHECK(hipEventCreate(&e));
            c->events.push_back(e);
        }
    }
    // with step_graph set, untimed runs of at least auto_steps steps replay a graph of that
    // many steps (launching the kernels one by one can cost more host time than the GPU
    // spends on a sparse step)
    const unsigned K = c->auto_steps;
    for (unsigned k = 0; k < nsteps;) {
        if (!time_it && K && nsteps - k >= K && c->drift_valid) {
            if (!c->agexec) {
                const int rc = capture_steps(c, K, &c->agraph, &c->agexec, c->apar);
                if (rc != XDEM_OK) return rc;
            }
            if (parities_match(c, c->apar)) {
                CTX_CHECK(hipGraphLaunch(c->agexec, c->stream));
                c->steps_done += K;
                k += K;
                continue;
            }
        }
        if (!time_it && K && c->drift_valid) {
            const int q = c->spp;
            if (!c->sgexec[q]) {
                const int rc = capture_steps(c, 1, &c->sgraph[q], &c->sgexec[q], c->spar[q]);
                if (rc != XDEM_OK) return rc;
            }
            if (parities_match(c, c->spar[q])) {
                CTX_CHECK(hipGraphLaunch(c->sgexec[q], c->stream));
                // the parities the step leaves (collide_half's flips)
                c->rbp ^= 1;
                c->set_pos_parity(c->spp ^ 1);
                c->cur ^= 1;
                if (c->cfg.fuse_velocity) c->set_state_parity(c->sv ^ 1);
                c->drift_valid = true;
                c->steps_done++;
                k++;
                continue;
            }
        }
        hipEvent_t* ev = (time_it && k < kmax) ? &c->events[kEventsPerStep * k] : nullptr;
        CTX_CHECK(c->full_step(ev));
        c->steps_done++;
        k++;
    }
    if (time_it) {
        CTX_CHECK(hipStreamSynchronize(c->stream));
        c->t_contact = 0.0;
        c->t_grid = 0.0;
        c->t_finish = 0.0;
        c->t_pairs = 0.0;
        c->t_lite_move = 0.0;
        c->t_collide = 0.0;
        c->t_pairpass = 0.0;
        c->t_meshpass = 0.0;
        c->t_steps = std::min(nsteps, kmax);
        for (unsigned k = 0; k < c->t_steps; k++) {
            const hipEvent_t* ev = &c->events[kEventsPerStep * k];
            float a = 0.f, p = 0.f, b = 0.f, f = 0.f, m = 0.f;
            CTX_CHECK(hipEventElapsedTime(&a, ev[0], ev[1]));
            CTX_CHECK(hipEventElapsedTime(&p, ev[1], ev[2]));
            CTX_CHECK(hipEventElapsedTime(&b, ev[2], ev[3]));
            CTX_CHECK(hipEventElapsedTime(&f, ev[3], ev[4]));
            CTX_CHECK(hipEventElapsedTime(&m, ev[5], ev[6]));
            float kc = 0.f, kp = 0.f, km = 0.f;
            CTX_CHECK(hipEventElapsedTime(&kc, ev[7], ev[8]));
            CTX_CHECK(hipEventElapsedTime(&km, ev[2], ev[9]));
            CTX_CHECK(hipEventElapsedTime(&kp, ev[9], ev[7]));
            c->t_collide += kc;
            c->t_meshpass += km;
            c->t_pairpass += kp;
            c->t_grid += a;
            c->t_pairs += p;
            c->t_contact += b;
            c->t_finish += f;
            c->t_lite_move += m;
        }
    }
    return XDEM_OK;
}

int xdem_ctx_step(xdem_ctx* c, unsigned int nsteps)
{
    if (!c) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    CTX_CHECK(hipSetDevice(c->cfg.device));
    return run_steps(c, nsteps);
}

int xdem_ctx_collide(xdem_ctx* c)
{
    if (!c) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    CTX_CHECK(hipSetDevice(c->cfg.device));
    CTX_CHECK(c->collide_half(false, nullptr));
    return XDEM_OK;
}

int xdem_ctx_finish(xdem_ctx* c)
{
    if (!c) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    CTX_CHECK(hipSetDevice(c->cfg.device));
    CTX_CHECK(c->finish_half(true));
    c->steps_done++;
    return XDEM_OK;
}

int xdem_ctx_synchronize(xdem_ctx* c)
{
    if (!c) return XDEM_E_INVALID;
    return hip_status(hipStreamSynchronize(c->stream));
}

static int field_info(xdem_ctx* c, int field, double** p, int* width)
{
    switch (field) {
    case XDEM_FIELD_POS: *p = c->pos; *width = 4; break;
    case XDEM_FIELD_VEL: *p = c->vel; *width = 3; break;
    case XDEM_FIELD_ACC: *p = c->acc; *width = 3; break;
    case XDEM_FIELD_EP: *p = c->ep; *width = 4; break;
    case XDEM_FIELD_EV: *p = c->ev; *width = 4; break;
    case XDEM_FIELD_EA: *p = c->ea; *width = 4; break;
    case XDEM_FIELD_FORCE: *p = c->force; *width = 3; break;
    case XDEM_FIELD_MOMENT: *p = c->moment; *width = 3; break;
    case XDEM_FIELD_MASS: *p = c->mass; *width = 1; break;
    case XDEM_FIELD_INERTIA: *p = c->inertia; *width = 3; break;
    case XDEM_FIELD_TMAX: *p = c->tmax; *width = 3; break;
    case XDEM_FIELD_RRES: *p = c->rres; *width = 1; break;
    default: return XDEM_E_INVALID;
    }
    return XDEM_OK;
}

int xdem_ctx_read(xdem_ctx* c, int field, unsigned int scene, double* out)
{
    if (!c || !out) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    if (scene >= c->scenes.size()) return XDEM_E_INVALID;
    double* p;
    int w;
    if (field_info(c, field, &p, &w)) return XDEM_E_INVALID;
    const SceneHost& h = c->scenes[scene];
    const size_t b = h.p_begin, m = h.s.np;
    std::vector<double> tmp((size_t)w * m);
    CTX_CHECK(hipSetDevice(c->cfg.device));
    CTX_CHECK(hipMemcpyAsync(tmp.data(), p + (size_t)w * b, sizeof(double) * w * m, hipMemcpyDeviceToHost, c->stream));
    CTX_CHECK(hipStreamSynchronize(c->stream));
    for (size_t j = 0; j < m; j++)   // stored row j is the caller's row ext_of - b
        std::copy(tmp.begin() + w * j, tmp.begin() + w * (j + 1), out + (size_t)w * (c->ext_of[b + j] - b));
    return XDEM_OK;
}

int xdem_ctx_write(xdem_ctx* c, int field, unsigned int scene, const double* in)
{
    if (!c || !in) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    if (scene >= c->scenes.size()) return XDEM_E_INVALID;
    double* p;
    int w;
    if (field_info(c, field, &p, &w)) return XDEM_E_INVALID;
    const SceneHost& h = c->scenes[scene];
    const size_t b = h.p_begin, m = h.s.np;
    std::vector<double> tmp((size_t)w * m);
    for (size_t j = 0; j < m; j++)
        std::copy(in + (size_t)w * (c->ext_of[b + j] - b), in + (size_t)w * (c->ext_of[b + j] - b + 1), tmp.begin() + w * j);
    CTX_CHECK(hipSetDevice(c->cfg.device));
    CTX_CHECK(hipMemcpyAsync(p + (size_t)w * b, tmp.data(), sizeof(double) * w * m, hipMemcpyHostToDevice, c->stream));
    CTX_CHECK(hipStreamSynchronize(c->stream));
    c->drift_valid = false;   // the precomputed drift used the old state
    return XDEM_OK;
}

int xdem_ctx_device_ptr(xdem_ctx* c, int field, void** out)
{
    if (!c || !out) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    double* p;
    int w;
    if (field_info(c, field, &p, &w)) return XDEM_E_INVALID;
    *out = p;
    c->drift_valid = false;   // the caller may write the state through the pointer
    return XDEM_OK;
}

// Re-sorts the stored order by the spheres' current positions: every per-sphere array that a
// step carries to the next is gathered into the new order (the sphere-pair slots' ids renamed),
// and the next step rebuilds the lists.
int xdem_ctx_reorder(xdem_ctx* c)
{
    if (!c) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    const size_t n = c->np;
    CTX_CHECK(hipSetDevice(c->cfg.device));
    hipStream_t s = c->stream;
    std::vector<double> pos(4 * n);
    CTX_CHECK(hipMemcpyAsync(pos.data(), c->pos, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, s));
    CTX_CHECK(hipStreamSynchronize(s));
    std::vector<uint32_t> perm(n), inv(n);
    for (const SceneHost& h : c->scenes) {
        const size_t b = h.p_begin, m = h.s.np;
        const Grid& g = h.dev.grid;
        std::vector<uint32_t> key(m), ord(m);
        std::iota(ord.begin(), ord.end(), 0u);
        for (size_t j = 0; j < m; j++) {
            const double* p = &pos[4 * (b + j)];
            const int gx = (int)std::floor((p[0] - g.wo_x) / g.cs);
            const int gy = (int)std::floor((p[1] - g.wo_y) / g.cs);
            const int gz = (int)std::floor((p[2] - g.wo_z) / g.cs);
            key[j] = ((uint32_t)gz & g.mask_z) * g.dim_xy + ((uint32_t)gy & g.mask_y) * g.dim_x + ((uint32_t)gx & g.mask_x);
        }
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return key[x] < key[y]; });
        for (size_t j = 0; j < m; j++) {
            perm[b + j] = (uint32_t)(b + ord[j]);
            inv[b + ord[j]] = (uint32_t)(b + j);
        }
    }
    // the caller's index of each stored sphere, and back
    std::vector<uint32_t> ext(n);
    for (size_t j = 0; j < n; j++) ext[j] = c->ext_of[perm[j]];
    c->ext_of.swap(ext);
    for (size_t j = 0; j < n; j++) c->int_of[c->ext_of[j]] = (uint32_t)j;

    // the widest row in 32-bit words: a state row (8), a kind's history slots (4 each) or the
    // wrench slots (12 per body)
    const size_t words = std::max({(size_t)8, 4u * (size_t)c->cap, 4u * (size_t)XDEM_MAX_P2PL_COUNT,
                                   4u * (size_t)XDEM_MAX_P2CY_COUNT, 4u * (size_t)XDEM_MAX_P2MS_COUNT,
                                   12u * (size_t)c->nwslot});
    uint32_t *d_perm = nullptr, *d_inv = nullptr, *tmp = nullptr;
    hipError_t e = hipMalloc(&d_perm, sizeof(uint32_t) * n);
    if (e == hipSuccess) e = hipMalloc(&d_inv, sizeof(uint32_t) * n);
    if (e == hipSuccess) e = hipMalloc(&tmp, sizeof(uint32_t) * words * n);
    if (e == hipSuccess) e = hipMemcpyAsync(d_perm, perm.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_inv, inv.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s);
    auto rows = [&](void* p, uint32_t w) {   // w 32-bit words per sphere
        if (e != hipSuccess || !p) return;
        const size_t t = n * w;
        hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s,
                           static_cast<const uint32_t*>(p), tmp, d_perm, n, w);
        e = hipMemcpyAsync(p, tmp, sizeof(uint32_t) * t, hipMemcpyDeviceToDevice, s);
    };
    rows(c->pos, 8);
    rows(c->ep, 8);
    if (c->drift_valid) {   // the next step's drifted positions
        rows(c->pos_buf[c->spp ^ 1], 8);
        rows(c->ep_buf[c->spp ^ 1], 8);
    }
    rows(c->vel, 6);
    rows(c->ev, 8);
    rows(c->acc, 6);
    rows(c->ea, 8);
    rows(c->force, 6);
    rows(c->moment, 6);
    rows(c->mass, 2);
    rows(c->inertia, 6);
    rows(c->tmax, 6);
    rows(c->rres, 2);
    rows(c->d_jflat, 2);
    rows(c->nbr_count, 1);
    rows(c->pp_count[c->cur], 1);
    rows(c->pp_tsd[c->cur], 4 * c->cap);
    if (e == hipSuccess) {   // the pair slots' ids, renamed below the (already gathered) count
        const size_t t = n * c->cap;
        hipLaunchKernelGGL(k_gather_pair_ids, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s, c->pp_id[c->cur], tmp,
                           c->pp_count[c->cur], d_perm, d_inv, n, c->cap);
        e = hipMemcpyAsync(c->pp_id[c->cur], tmp, sizeof(uint32_t) * t, hipMemcpyDeviceToDevice, s);
    }
    rows(c->pl_count, 1);
    rows(c->pl_id, XDEM_MAX_P2PL_COUNT);
    rows(c->pl_tsd, 4 * XDEM_MAX_P2PL_COUNT);
    if (!c->cyls.empty()) {
        rows(c->cy_count, 1);
        rows(c->cy_id, XDEM_MAX_P2CY_COUNT);
        rows(c->cy_tsd, 4 * XDEM_MAX_P2CY_COUNT);
    }
    if (!c->meshes.empty()) {
        rows(c->ms_count, 1);
        rows(c->ms_id, XDEM_MAX_P2MS_COUNT);
        rows(c->ms_tsd, 4 * XDEM_MAX_P2MS_COUNT);
    }
    if (e == hipSuccess && c->nwslot) {   // the wrench slots (columns of n)
        const size_t t = n * 6 * c->nwslot;
        hipLaunchKernelGGL(k_gather_cols, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s, c->wrench_pp,
                           reinterpret_cast<double*>(tmp), d_perm, n, 6u * c->nwslot);
        e = hipMemcpyAsync(c->wrench_pp, tmp, sizeof(double) * t, hipMemcpyDeviceToDevice, s);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_ext, c->ext_of.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_int, c->int_of.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {   // the next step builds the lists (it reads the flag of its parity)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, c->rebuild + (c->rbp ^ 1), 1u, (size_t)1);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(tmp);
    (void)hipFree(d_inv);
    (void)hipFree(d_perm);
    return hip_status(e);
}

int xdem_ctx_sphere_order(xdem_ctx* c, unsigned int scene, unsigned int* order)
{
    if (!c || !order) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    if (scene >= c->scenes.size()) return XDEM_E_INVALID;
    const SceneHost& h = c->scenes[scene];
    for (size_t j = 0; j < h.s.np; j++) order[j] = c->ext_of[h.p_begin + j] - h.p_begin;
    return XDEM_OK;
}

int xdem_ctx_read_history(xdem_ctx* c, int kind, unsigned int scene, unsigned int* count, unsigned int* ids, double* tsd)
{
    if (!c || !count || !ids || !tsd) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    if (scene >= c->scenes.size()) return XDEM_E_INVALID;
    const SceneHost& h = c->scenes[scene];
    const size_t b = h.p_begin, m = h.s.np;
    CTX_CHECK(hipSetDevice(c->cfg.device));
    uint32_t cap;
    const uint32_t *dc, *di;
    const double2* dt;
    if (kind == XDEM_HISTORY_P2P) {
        cap = c->cap; dc = c->pp_count[c->cur]; di = c->pp_id[c->cur]; dt = c->pp_tsd[c->cur];
    } else if (kind == XDEM_HISTORY_PLANE) {
        cap = XDEM_MAX_P2PL_COUNT; dc = c->pl_count; di = c->pl_id; dt = c->pl_tsd;
    } else if (kind == XDEM_HISTORY_MESH) {
        if (h.dev.mesh < 0) return XDEM_E_INVALID;
        cap = XDEM_MAX_P2MS_COUNT; dc = c->ms_count; di = c->ms_id; dt = c->ms_tsd;
    } else if (kind == XDEM_HISTORY_CYLINDER) {
        if (h.dev.ncyl == 0) return XDEM_E_INVALID;
        cap = XDEM_MAX_P2CY_COUNT; dc = c->cy_count; di = c->cy_id; dt = c->cy_tsd;
    } else {
        return XDEM_E_INVALID;
    }
    std::vector<uint32_t> tc(m), ti(m * cap);
    std::vector<double> tt(2 * m * cap);
    CTX_CHECK(hipMemcpyAsync(tc.data(), dc + b, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, c->stream));
    CTX_CHECK(hipMemcpyAsync(ti.data(), di + b * cap, sizeof(uint32_t) * m * cap, hipMemcpyDeviceToHost, c->stream));
    CTX_CHECK(hipMemcpyAsync(tt.data(), dt + b * cap, sizeof(double2) * m * cap, hipMemcpyDeviceToHost, c->stream));
    CTX_CHECK(hipStreamSynchronize(c->stream));
    // stored row j is the caller's row ext_of - b; a sphere-pair slot holds a stored sphere,
    // returned as the caller's scene-local index. The sphere-pair slots past the count (left
    // over from earlier steps in the step's double buffer) read as zeros.
    for (size_t j = 0; j < m; j++) {
        const size_t e = c->ext_of[b + j] - b;
        count[e] = tc[j];
        for (uint32_t q = 0; q < cap; q++) {
            const bool unused = kind == XDEM_HISTORY_P2P && q >= tc[j];
            uint32_t id = ti[j * cap + q];
            if (kind == XDEM_HISTORY_P2P) id = unused ? 0u : c->ext_of[id] - (uint32_t)b;
            ids[e * cap + q] = id;
            tsd[2 * (e * cap + q)] = unused ? 0.0 : tt[2 * (j * cap + q)];
            tsd[2 * (e * cap + q) + 1] = unused ? 0.0 : tt[2 * (j * cap + q) + 1];
        }
    }
    return XDEM_OK;
}

int xdem_ctx_write_history(xdem_ctx* c, int kind, unsigned int scene, const unsigned int* count,
                           const unsigned int* ids, const double* tsd)
{
    if (!c || !count || !ids || !tsd) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    if (scene >= c->scenes.size()) return XDEM_E_INVALID;
    const SceneHost& h = c->scenes[scene];
    const size_t b = h.p_begin, m = h.s.np;
    uint32_t cap, id_limit;
    uint32_t *dc, *di;
    double2* dt;
    if (kind == XDEM_HISTORY_P2P) {
        cap = c->cap; dc = c->pp_count[c->cur]; di = c->pp_id[c->cur]; dt = c->pp_tsd[c->cur];
        id_limit = (uint32_t)m;
    } else if (kind == XDEM_HISTORY_PLANE) {
        cap = XDEM_MAX_P2PL_COUNT; dc = c->pl_count; di = c->pl_id; dt = c->pl_tsd;
        id_limit = UINT32_MAX;
    } else if (kind == XDEM_HISTORY_MESH) {
        if (h.dev.mesh < 0) return XDEM_E_INVALID;
        cap = XDEM_MAX_P2MS_COUNT; dc = c->ms_count; di = c->ms_id; dt = c->ms_tsd;
        id_limit = c->meshes[h.dev.mesh].dev.ntri;
    } else if (kind == XDEM_HISTORY_CYLINDER) {
        if (h.dev.ncyl == 0) return XDEM_E_INVALID;
        cap = XDEM_MAX_P2CY_COUNT; dc = c->cy_count; di = c->cy_id; dt = c->cy_tsd;
        id_limit = UINT32_MAX;   // cylinder ids (and id + 1000) index nothing
    } else {
        return XDEM_E_INVALID;
    }
    // a count past the slots or an id outside the scene would index past its arrays in the
    // contact kernels: refuse the whole history
    for (size_t e = 0; e < m; e++) {
        if (count[e] > cap) return XDEM_E_INVALID;
        for (uint32_t q = 0; q < count[e]; q++)
            if (id_limit != UINT32_MAX && ids[e * cap + q] >= id_limit) return XDEM_E_INVALID;
    }
    // the caller's row e goes to stored row int_of - b; a sphere-pair slot's caller's index to
    // its stored sphere
    std::vector<uint32_t> tc(m), ti(m * cap);
    std::vector<double> tt(2 * m * cap);
    for (size_t e = 0; e < m; e++) {
        const size_t j = c->int_of[b + e] - b;
        tc[j] = count[e];
        for (uint32_t q = 0; q < cap; q++) {
            uint32_t id = ids[e * cap + q];
            if (kind == XDEM_HISTORY_P2P && q < count[e]) id = c->int_of[b + id];
            ti[j * cap + q] = id;
            tt[2 * (j * cap + q)] = tsd[2 * (e * cap + q)];
            tt[2 * (j * cap + q) + 1] = tsd[2 * (e * cap + q) + 1];
        }
    }
    CTX_CHECK(hipSetDevice(c->cfg.device));
    CTX_CHECK(hipMemcpyAsync(dc + b, tc.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice, c->stream));
    CTX_CHECK(hipMemcpyAsync(di + b * cap, ti.data(), sizeof(uint32_t) * m * cap, hipMemcpyHostToDevice, c->stream));
    CTX_CHECK(hipMemcpyAsync(dt + b * cap, tt.data(), sizeof(double2) * m * cap, hipMemcpyHostToDevice, c->stream));
    CTX_CHECK(hipStreamSynchronize(c->stream));
    return XDEM_OK;
}

int xdem_ctx_scene_info_get(xdem_ctx* c, unsigned int scene, xdem_ctx_scene_info* out)
{
    if (!c || !out) return XDEM_E_INVALID;
    if (scene >= c->scenes.size()) return XDEM_E_INVALID;
    const SceneHost& h = c->scenes[scene];
    std::memset(out, 0, sizeof(*out));
    out->np = h.s.np;
    out->nplane = h.s.nplane;
    out->p2p_capacity = c->cap;
    out->has_mesh = h.dev.mesh >= 0 ? 1u : 0u;
    out->first_particle = h.p_begin;
    for (const CylHost& y : c->cyls) out->ncylinder += y.scene == scene ? 1u : 0u;
    return XDEM_OK;
}

int xdem_ctx_error_word(xdem_ctx* c, unsigned int* out, int clear)
{
    if (!c || !out) return XDEM_E_INVALID;
    if (!c->committed) { *out = 0; return XDEM_OK; }
    CTX_CHECK(hipSetDevice(c->cfg.device));
    CTX_CHECK(hipMemcpyAsync(out, c->err, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    CTX_CHECK(hipStreamSynchronize(c->stream));
    if (clear) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, c->err, 0u, (size_t)1);
        CTX_CHECK(hipGetLastError());
        CTX_CHECK(hipStreamSynchronize(c->stream));
    }
    return XDEM_OK;
}

int xdem_ctx_body_wrench(xdem_ctx* c, unsigned int scene, double* out)
{
    if (!c || !out) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    if (scene >= c->scenes.size()) return XDEM_E_INVALID;
    const SceneHost& h = c->scenes[scene];
    std::vector<double> all(6 * (size_t)c->nwslot);
    CTX_CHECK(hipSetDevice(c->cfg.device));
    if (c->nwslot) {
        CTX_CHECK(hipMemcpyAsync(all.data(), c->wrench_out, sizeof(double) * 6 * c->nwslot, hipMemcpyDeviceToHost, c->stream));
        CTX_CHECK(hipStreamSynchronize(c->stream));
    }
    for (unsigned k = 0; k < h.s.nplane; k++) {
        const int slot = c->wall_slot[h.wall_begin + k];
        for (int q = 0; q < 6; q++) out[6 * k + q] = slot >= 0 ? all[6 * (size_t)slot + q] : 0.0;
    }
    return XDEM_OK;
}

int xdem_ctx_set_body_pose(xdem_ctx* c, unsigned int scene, unsigned int wall, const double dbi[15])
{
    if (!c || !dbi) return XDEM_E_INVALID;
    if (scene >= c->scenes.size()) return XDEM_E_INVALID;
    SceneHost& h = c->scenes[scene];
    if (wall >= h.s.nplane) return XDEM_E_INVALID;
    WallDev& W = c->walls[h.wall_begin + wall];
    // xParticlePlaneContact::update (xParticlePlaneContact.cpp:96-110): corner k = pose
    // position + A(e) LocalPoint(k), with the local corners fixed at add_scene
    double g[4][3];
    for (int q = 0; q < 4; q++) {
        double gl[3];
        euler_rotate(dbi + XDEM_E0, &h.wall_local[12 * wall + 3 * q], false, gl);
        for (int a = 0; a < 3; a++) g[q][a] = dbi[XDEM_PX + a] + gl[a];
    }
    xdem_plane_info pi = W.info;
    pi.xw = {g[0][0], g[0][1], g[0][2]};
    pi.w2 = {g[1][0], g[1][1], g[1][2]};
    pi.w3 = {g[2][0], g[2][1], g[2][2]};
    pi.w4 = {g[3][0], g[3][1], g[3][2]};
    const double pa[3] = {g[1][0] - g[0][0], g[1][1] - g[0][1], g[1][2] - g[0][2]};
    const double pb[3] = {g[3][0] - g[0][0], g[3][1] - g[0][1], g[3][2] - g[0][2]};
    pi.pa = {pa[0], pa[1], pa[2]};
    pi.pb = {pb[0], pb[1], pb[2]};
    pi.l1 = std::sqrt(pa[0] * pa[0] + pa[1] * pa[1] + pa[2] * pa[2]);
    pi.l2 = std::sqrt(pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2]);
    // the largest displacement of the wall's points is that of a corner (rigid motion)
    double dmax = 0.0;
    {
        const xdem_double3 oc[4] = {W.info.xw, W.info.w2, W.info.w3, W.info.w4};
        for (int q = 0; q < 4; q++) {
            const double dx = g[q][0] - oc[q].x, dy = g[q][1] - oc[q].y, dz = g[q][2] - oc[q].z;
            dmax = std::max(dmax, std::sqrt(dx * dx + dy * dy + dz * dz));
        }
        dmax = dmax * (1.0 + 1e-9) + 1e-15;   // rounded up
    }
    const double u1[3] = {pa[0] / pi.l1, pa[1] / pi.l1, pa[2] / pi.l1};
    const double u2[3] = {pb[0] / pi.l2, pb[1] / pi.l2, pb[2] / pi.l2};
    pi.u1 = {u1[0], u1[1], u1[2]};
    pi.u2 = {u2[0], u2[1], u2[2]};
    pi.uw = {u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2], u1[0] * u2[1] - u1[1] * u2[0]};
    W.info = pi;
    std::memcpy(W.dbi, dbi, sizeof(double) * 15);
    if (c->committed) {
        CTX_CHECK(hipSetDevice(c->cfg.device));
        CTX_CHECK(hipMemcpyAsync(c->d_walls + h.wall_begin + wall, &W, sizeof(WallDev), hipMemcpyHostToDevice, c->stream));
        CTX_CHECK(c->body_moved(dmax));   // the lists' wall-reach flags hold while it moved less than the skin
        CTX_CHECK(hipStreamSynchronize(c->stream));
    }
    return XDEM_OK;
}

int xdem_ctx_add_mesh(xdem_ctx* c, unsigned int scene, const xdem_mesh* m, unsigned int* idx)
{
    if (!c || !m || !m->vertices || m->ntriangle == 0) return XDEM_E_INVALID;
    if (c->committed) return XDEM_E_NOT_READY;
    if (scene >= c->scenes.size()) return XDEM_E_INVALID;
    SceneHost& sh = c->scenes[scene];
    if (sh.dev.mesh >= 0) return XDEM_E_UNSUPPORTED;   // one mesh body per scene
    if ((unsigned long long)c->ntri + m->ntriangle >= (1ull << kIdBits)) return XDEM_E_INVALID;
    const uint32_t T = m->ntriangle;
    MeshHost h;
    h.scene = scene;
    h.vlist.assign(m->vertices, m->vertices + 9 * (size_t)T);
    for (double v : h.vlist)
        if (!std::isfinite(v)) return XDEM_E_INVALID;
    const double* d = m->dbi;
    // proxy spheres: fitted to the global triangles at this pose, kept in the body frame
    // (xParticleMeshObjectContact.cpp:104-127)
    h.plocal.resize(3 * (size_t)T);
    h.prad.resize(T);
    double maxprox = 0.0;
    for (uint32_t t = 0; t < T; t++) {
        double g[3][3];
        for (int q = 0; q < 3; q++) {
            double gl[3];
            euler_rotate(d + XDEM_E0, &h.vlist[9 * (size_t)t + 3 * q], false, gl);
            for (int a = 0; a < 3; a++) g[q][a] = d[XDEM_PX + a] + gl[a];
        }
        double sph[4];
        xdem_fit_sphere_to_triangle(g[0], g[1], g[2], 0.8, sph);
        const double rel[3] = {sph[0] - d[XDEM_PX], sph[1] - d[XDEM_PY], sph[2] - d[XDEM_PZ]};
        euler_rotate(d + XDEM_E0, rel, true, &h.plocal[3 * (size_t)t]);
        h.prad[t] = sph[3];
        maxprox = std::max(maxprox, sph[3]);
        for (int q = 0; q < 3; q++) {
            const double* v = &h.vlist[9 * (size_t)t + 3 * q];
            h.rmax_local = std::max(h.rmax_local, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
        }
    }
    MeshDev& D = h.dev;
    std::memset(&D, 0, sizeof(D));
    std::memcpy(D.dbi, d, sizeof(double) * 15);
    D.cc = make_contact_const(m->cp, sh.rmax);
    D.tri_begin = c->ntri;
    D.ntri = T;
    D.wslot = -1;
    // body-frame grid: every triangle in each cell its bounding box, grown by the largest
    // reach of a list (radius, cohesion reach and skin at scale 1, with a rounding margin),
    // overlaps; cells of twice that reach or the mean edge length, at most 2^22 cells
    const double Rq = (sh.rmax + D.cc.coh_reach + sh.dev.skin) * (1.0 + 1e-6) + 1e-9;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, edges = 0.0;
    for (uint32_t t = 0; t < T; t++)
        for (int q = 0; q < 3; q++) {
            const double* v = &h.vlist[9 * (size_t)t + 3 * q];
            const double* w = &h.vlist[9 * (size_t)t + 3 * ((q + 1) % 3)];
            edges += std::sqrt((v[0] - w[0]) * (v[0] - w[0]) + (v[1] - w[1]) * (v[1] - w[1]) + (v[2] - w[2]) * (v[2] - w[2]));
            for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], v[a]); hi[a] = std::max(hi[a], v[a]); }
        }
    for (int a = 0; a < 3; a++) { lo[a] -= Rq; hi[a] += Rq; }
    double hc = std::max(2.0 * Rq, edges / (3.0 * T));
    uint32_t dims[3];
    for (;;) {
        unsigned long long cells = 1;
        for (int a = 0; a < 3; a++) {
            dims[a] = (uint32_t)std::max(1.0, std::ceil((hi[a] - lo[a]) / hc));
            cells *= dims[a];
        }
        if (cells <= (1ull << 22)) break;
        hc *= 1.25;
    }
    const size_t ncell = (size_t)dims[0] * dims[1] * dims[2];
    auto cell_range = [&](uint32_t t, int a, uint32_t& c0, uint32_t& c1) {
        double mn = 1e300, mx = -1e300;
        for (int q = 0; q < 3; q++) {
            mn = std::min(mn, h.vlist[9 * (size_t)t + 3 * q + a]);
            mx = std::max(mx, h.vlist[9 * (size_t)t + 3 * q + a]);
        }
        const double f0 = std::floor((mn - Rq - lo[a]) / hc), f1 = std::floor((mx + Rq - lo[a]) / hc);
        c0 = (uint32_t)std::max(0.0, std::min(f0, (double)dims[a] - 1.0));
        c1 = (uint32_t)std::max(0.0, std::min(f1, (double)dims[a] - 1.0));
    };
    h.cell_start.assign(ncell + 1, 0u);
    for (int pass = 0; pass < 2; pass++) {
        std::vector<uint32_t> fill;
        if (pass == 1) {
            for (size_t k = 0; k < ncell; k++) h.cell_start[k + 1] += h.cell_start[k];
            h.cell_tris.resize(h.cell_start[ncell]);
            fill.assign(h.cell_start.begin(), h.cell_start.end() - 1);
        }
        for (uint32_t t = 0; t < T; t++) {
            uint32_t r[3][2];
            for (int a = 0; a < 3; a++) cell_range(t, a, r[a][0], r[a][1]);
            for (uint32_t z = r[2][0]; z <= r[2][1]; z++)
                for (uint32_t y = r[1][0]; y <= r[1][1]; y++)
                    for (uint32_t x = r[0][0]; x <= r[0][1]; x++) {
                        const size_t cell = ((size_t)z * dims[1] + y) * dims[0] + x;
                        if (pass == 0) h.cell_start[cell + 1]++;
                        else h.cell_tris[fill[cell]++] = t;
                    }
        }
    }
    D.lo_x = lo[0]; D.lo_y = lo[1]; D.lo_z = lo[2]; D.h = hc;
    D.dim_x = dims[0]; D.dim_y = dims[1]; D.dim_z = dims[2];
    if (d[XDEM_MM] > 0.0) {   // a dynamic body: its wrench is reduced every step
        D.wslot = (int32_t)c->nwslot++;
        c->wslots.push_back({sh.dev.p_begin, sh.dev.p_end, kMeshNear, 0u});
    }
    // the reference's cell is twice the largest of the sphere and proxy radii
    sh.dev.ref_cs = std::max(sh.dev.ref_cs, 2.0 * maxprox);
    sh.dev.mesh = (int32_t)c->meshes.size();
    c->ntri += T;
    if (idx) *idx = (unsigned)c->meshes.size();
    c->meshes.push_back(std::move(h));
    return XDEM_OK;
}

int xdem_ctx_set_mesh_pose(xdem_ctx* c, unsigned int mesh, const double dbi[15])
{
    if (!c || !dbi) return XDEM_E_INVALID;
    if (mesh >= c->meshes.size()) return XDEM_E_INVALID;
    for (int q = 0; q < 15; q++)
        if (!std::isfinite(dbi[q])) return XDEM_E_INVALID;
    MeshHost& m = c->meshes[mesh];
    // largest displacement of a point of the body: |dP| + |A(e') - A(e)|_F r_max, rounded up
    double A0[9], A1[9];
    for (int k = 0; k < 3; k++) {
        const double ek[3] = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
        double c0[3], c1[3];
        euler_rotate(m.dev.dbi + XDEM_E0, ek, false, c0);
        euler_rotate(dbi + XDEM_E0, ek, false, c1);
        for (int a = 0; a < 3; a++) { A0[3 * a + k] = c0[a]; A1[3 * a + k] = c1[a]; }
    }
    double fro = 0.0, dp = 0.0;
    for (int k = 0; k < 9; k++) fro += (A1[k] - A0[k]) * (A1[k] - A0[k]);
    for (int a = 0; a < 3; a++) dp += (dbi[XDEM_PX + a] - m.dev.dbi[XDEM_PX + a]) * (dbi[XDEM_PX + a] - m.dev.dbi[XDEM_PX + a]);
    const double dmax = (std::sqrt(dp) + std::sqrt(fro) * m.rmax_local) * (1.0 + 1e-9) + 1e-15;
    std::memcpy(m.dev.dbi, dbi, sizeof(double) * 15);
    if (c->committed) {
        CTX_CHECK(hipSetDevice(c->cfg.device));
        PoseArg pa;
        std::memcpy(pa.v, dbi, sizeof(pa.v));
        hipLaunchKernelGGL(k_set_mesh_pose, dim3(1), dim3(64), 0, c->stream, c->d_meshes, mesh, pa);
        hipLaunchKernelGGL(k_mesh_pose, dim3((m.dev.ntri + 255) / 256), dim3(256), 0, c->stream, c->d_meshes, mesh,
                           c->d_vlist, c->d_plocal, c->d_prad, c->d_tri, c->d_prox);
        CTX_CHECK(hipGetLastError());
        CTX_CHECK(c->body_moved(dmax));
    }
    return XDEM_OK;
}

int xdem_ctx_mesh_wrench(xdem_ctx* c, unsigned int mesh, double* out)
{
    if (!c || !out) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    if (mesh >= c->meshes.size()) return XDEM_E_INVALID;
    const int slot = c->meshes[mesh].dev.wslot;
    for (int q = 0; q < 6; q++) out[q] = 0.0;
    if (slot < 0) return XDEM_OK;
    CTX_CHECK(hipSetDevice(c->cfg.device));
    CTX_CHECK(hipMemcpyAsync(c->h_wrench, c->wrench_out + 6 * (size_t)slot, sizeof(double) * 6, hipMemcpyDeviceToHost,
                             c->stream));
    CTX_CHECK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < 6; q++) out[q] = c->h_wrench[q];
    return XDEM_OK;
}

// Margin of a cylinder's reach test. In each contact region the overlap is r plus a function of
// the centre that changes by at most the distance the centre moves; where two regions meet the
// functions differ by at most: 2 |side radius - rim radius| (outside side wall / rim, inside /
// outside of the tube), the lid ring's thickness and, with one, the rim radius (a centre
// beyond an end inside the rim radius sees the rim, inside the tube the lid), twice the lid
// planes' offset from the axis ends (body position off the axis midpoint, length other than
// the axis) and, for a cylinder not longer than a sphere's reach, its length (the nearer rim is
// taken by the local y offset against the length). The geometry is that of a cylinder about
// the body frame's y axis only when both axis ends lie on that axis and are a length apart:
// any other is always in reach (margin 1e300).
static double cylinder_margin(const xdem_cylinder_info& y, double rmax, double skin)
{
    const double len = y.len_rr.x;
    const double ab[3] = {y.ptop.x - y.pbase.x, y.ptop.y - y.pbase.y, y.ptop.z - y.pbase.z};
    const double lab = std::sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
    const double tol = 1e-9 * std::max(len, lab);
    const bool on_axis = std::fabs(y.pbase.x) <= tol && std::fabs(y.pbase.z) <= tol && std::fabs(y.ptop.x) <= tol &&
                         std::fabs(y.ptop.z) <= tol;
    if (!on_axis || std::fabs(lab - len) > tol) return 1e300;
    const double mid = std::fabs(0.5 * (y.pbase.y + y.ptop.y));
    double m = 2.0 * std::fabs(y.len_rr.y - y.len_rr.z) + y.thickness + (y.thickness > 0.0 ? y.len_rr.z : 0.0) +
               std::fabs(len - lab) + 2.0 * mid;
    if (len <= rmax + skin) m += len;
    return m * (1.0 + 1e-9) + tol;
}

int xdem_ctx_add_cylinder(xdem_ctx* c, unsigned int scene, const xdem_cylinder_info* cy, const double dbi[15],
                          const xdem_contact_property* cp, unsigned int* idx)
{
    if (!c || !cy || !dbi || !cp) return XDEM_E_INVALID;
    if (c->committed) return XDEM_E_NOT_READY;
    if (scene >= c->scenes.size()) return XDEM_E_INVALID;
    for (int q = 0; q < 15; q++)
        if (!std::isfinite(dbi[q])) return XDEM_E_INVALID;
    const xdem_cylinder_info& y = *cy;
    const double g[10] = {y.thickness, y.len_rr.x, y.len_rr.y, y.len_rr.z, y.pbase.x,
                          y.pbase.y,   y.pbase.z,  y.ptop.x,   y.ptop.y,   y.ptop.z};
    for (double v : g)
        if (!std::isfinite(v)) return XDEM_E_INVALID;
    if (!(y.len_rr.x > 0.0) || !(y.len_rr.y >= 0.0) || !(y.len_rr.z > 0.0) || !(y.thickness >= 0.0)) return XDEM_E_INVALID;
    if (y.empty_part > 2u || y.id > UINT32_MAX - 1000u) return XDEM_E_INVALID;   // id + 1000 names the inner lids
    if (y.ptop.x == y.pbase.x && y.ptop.y == y.pbase.y && y.ptop.z == y.pbase.z) return XDEM_E_INVALID;
    unsigned nin = 0;
    for (const CylHost& o : c->cyls) nin += o.scene == scene ? 1u : 0u;
    if (nin >= (unsigned)kMaxWallsPerScene) return XDEM_E_INVALID;
    const SceneHost& sh = c->scenes[scene];
    CylHost h;
    h.scene = scene;
    CylDev& D = h.dev;
    std::memset(&D, 0, sizeof(D));
    D.info = y;
    std::memcpy(D.dbi, dbi, sizeof(double) * 15);
    D.cc = make_contact_const(*cp, sh.rmax);
    if (cp->rfric == 0.0) c->tmax_persist = 1;
    D.margin = cylinder_margin(y, sh.rmax, sh.dev.skin);
    D.wslot = -1;
    if (dbi[XDEM_MM] > 0.0) {   // a cylinder with mass is a dynamic body: reduce its wrench
        D.wslot = (int32_t)c->nwslot++;
        c->wslots.push_back({sh.dev.p_begin, sh.dev.p_end, kTouch, 0u});
    }
    // body-frame radius of the points whose overlap can change with the pose: the axis ends,
    // the side wall or the lid ring, and a sphere's reach beyond them
    const double eb = std::sqrt(y.pbase.x * y.pbase.x + y.pbase.y * y.pbase.y + y.pbase.z * y.pbase.z);
    const double et = std::sqrt(y.ptop.x * y.ptop.x + y.ptop.y * y.ptop.y + y.ptop.z * y.ptop.z);
    h.extent = std::max(eb, et) + std::max(y.len_rr.y, y.len_rr.z + y.thickness) + sh.rmax + sh.dev.skin +
               (D.margin < 1e300 ? D.margin : 0.0);
    if (idx) *idx = (unsigned)c->cyls.size();
    c->cyls.push_back(h);
    return XDEM_OK;
}

int xdem_ctx_set_cylinder_pose(xdem_ctx* c, unsigned int cyl, const double dbi[15])
{
    if (!c || !dbi) return XDEM_E_INVALID;
    if (cyl >= c->cyls.size()) return XDEM_E_INVALID;
    for (int q = 0; q < 15; q++)
        if (!std::isfinite(dbi[q])) return XDEM_E_INVALID;
    CylHost& h = c->cyls[cyl];
    // The geometry is unchanged by a rotation about the cylinder's axis (the body frame's y
    // axis): a point of it moves by at most |dP| + |a' - a| extent, a = A(e) (0, 1, 0) the
    // axis direction (the least rotation taking a to a' moves a unit vector by |a' - a|).
    // Spinning a drum about its axis does not wear the neighbour lists out.
    double a0[3], a1[3];
    const double ey[3] = {0.0, 1.0, 0.0};
    euler_rotate(h.dev.dbi + XDEM_E0, ey, false, a0);
    euler_rotate(dbi + XDEM_E0, ey, false, a1);
    double da = 0.0, dp = 0.0;
    for (int k = 0; k < 3; k++) {
        da += (a1[k] - a0[k]) * (a1[k] - a0[k]);
        dp += (dbi[XDEM_PX + k] - h.dev.dbi[XDEM_PX + k]) * (dbi[XDEM_PX + k] - h.dev.dbi[XDEM_PX + k]);
    }
    const double dmax = (std::sqrt(dp) + std::sqrt(da) * h.extent) * (1.0 + 1e-9) + 1e-15;
    std::memcpy(h.dev.dbi, dbi, sizeof(double) * 15);
    if (c->committed) {
        CTX_CHECK(hipSetDevice(c->cfg.device));
        CTX_CHECK(hipMemcpyAsync(c->d_cyls + c->cyl_dev[cyl], &h.dev, sizeof(CylDev), hipMemcpyHostToDevice, c->stream));
        if (h.dev.margin < 1e300) CTX_CHECK(c->body_moved(dmax));   // (a cylinder always in reach needs no lists)
        CTX_CHECK(hipStreamSynchronize(c->stream));   // the host's pose copy may change after return
    }
    return XDEM_OK;
}

int xdem_ctx_cylinder_wrench(xdem_ctx* c, unsigned int cyl, double* out)
{
    if (!c || !out) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    if (cyl >= c->cyls.size()) return XDEM_E_INVALID;
    const int slot = c->cyls[cyl].dev.wslot;
    for (int q = 0; q < 6; q++) out[q] = 0.0;
    if (slot < 0) return XDEM_OK;
    CTX_CHECK(hipSetDevice(c->cfg.device));
    CTX_CHECK(hipMemcpyAsync(c->h_wrench, c->wrench_out + 6 * (size_t)slot, sizeof(double) * 6, hipMemcpyDeviceToHost,
                             c->stream));
    CTX_CHECK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < 6; q++) out[q] = c->h_wrench[q];
    return XDEM_OK;
}

int xdem_ctx_info_get(xdem_ctx* c, xdem_ctx_info* out)
{
    if (!c || !out) return XDEM_E_INVALID;
    std::memset(out, 0, sizeof(*out));
    out->nscene = (unsigned)c->scenes.size();
    out->np = c->np;
    out->nsorted = c->np;
    out->nplane_total = (unsigned)c->walls.size();
    out->nbucket = c->nbucket;
    out->key_bits = c->key_bits;
    out->p2p_capacity = c->cap;
    out->device_bytes = c->bytes;
    out->ntriangle = c->ntri;
    out->steps_done = c->steps_done;
    if (c->committed) {
        CTX_CHECK(hipSetDevice(c->cfg.device));
        CTX_CHECK(hipMemcpyAsync(&out->list_builds, c->n_rebuild, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                 c->stream));
        std::vector<uint32_t> q1(kQueueSegs * kSegStride), q2(kQueueSegs * kSegStride);
        CTX_CHECK(hipMemcpyAsync(q1.data(), c->seg_count, sizeof(uint32_t) * q1.size(), hipMemcpyDeviceToHost,
                                 c->stream));
        CTX_CHECK(hipMemcpyAsync(q2.data(), c->seg_count2, sizeof(uint32_t) * q2.size(), hipMemcpyDeviceToHost,
                                 c->stream));
        CTX_CHECK(hipStreamSynchronize(c->stream));
        for (uint32_t k = 0; k < kQueueSegs; k++) {
            out->list_spheres += q1[k * kSegStride];
            out->contact_spheres += q2[k * kSegStride];
        }
        if (c->mesh_count) {
            CTX_CHECK(hipMemcpyAsync(&out->mesh_spheres, c->mesh_count, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                     c->stream));
            CTX_CHECK(hipStreamSynchronize(c->stream));
        }
    }
    return XDEM_OK;
}

int xdem_ctx_contact_count(xdem_ctx* c, unsigned long long* n_p2p, unsigned long long* n_wall)
{
    if (!c || !n_p2p || !n_wall) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    unsigned long long h[2];
    CTX_CHECK(hipSetDevice(c->cfg.device));
    CTX_CHECK(hipMemcpyAsync(h, c->counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    CTX_CHECK(hipStreamSynchronize(c->stream));
    *n_p2p = h[0];
    *n_wall = h[1];
    return XDEM_OK;
}

int xdem_ctx_stream(xdem_ctx* c, void** out)
{
    if (!c || !out) return XDEM_E_INVALID;
    *out = c->stream;
    return XDEM_OK;
}

int xdem_ctx_graph_capture(xdem_ctx* c, unsigned int nsteps)
{
    if (!c) return XDEM_E_INVALID;
    if (!c->committed) return XDEM_E_NOT_READY;
    if (nsteps == 0 || (nsteps & 1u)) return XDEM_E_INVALID;  // even: history buffers return to the same parity
    CTX_CHECK(hipSetDevice(c->cfg.device));
    // the old graph is replaced; destroying it cannot fail in a way the caller could act on
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    int par[4];
    const int rc = capture_steps(c, nsteps, &c->graph, &c->gexec, par);
    if (rc != XDEM_OK) return rc;
    c->graph_parity = par[0];
    c->graph_sparity = par[1];
    c->graph_pparity = par[2];
    c->graph_steps = nsteps;
    return XDEM_OK;
}

int xdem_ctx_graph_launch(xdem_ctx* c, unsigned int nlaunch)
{
    if (!c) return XDEM_E_INVALID;
    if (!c->gexec) return XDEM_E_NOT_READY;
    // the graph's buffers are those of the parity it was recorded at (an odd number of
    // eager steps since then swaps them)
    if (c->cur != c->graph_parity || c->sv != c->graph_sparity || c->spp != c->graph_pparity || !c->drift_valid)
        return XDEM_E_NOT_READY;
    CTX_CHECK(hipSetDevice(c->cfg.device));
    for (unsigned k = 0; k < nlaunch; k++) {
        CTX_CHECK(hipGraphLaunch(c->gexec, c->stream));
        c->steps_done += c->graph_steps;
    }
    return XDEM_OK;
}

int xdem_ctx_set_timing(xdem_ctx* c, int on)
{
    if (!c) return XDEM_E_INVALID;
    c->timing = on != 0;
    return XDEM_OK;
}

int xdem_ctx_timing(xdem_ctx* c, double out[9])
{
    if (!c || !out) return XDEM_E_INVALID;
    out[6] = c->t_collide;
    out[7] = c->t_meshpass;
    out[8] = c->t_pairpass;
    out[5] = c->t_lite_move;
    out[0] = c->t_contact;
    out[1] = c->t_grid;
    out[2] = (double)c->t_steps;
    out[3] = c->t_finish;
    out[4] = c->t_pairs;
    return XDEM_OK;
}

}  // extern "C"

// xdem_launch.hip -- launcher-level C ABI: one gfx950 kernel per launcher of the
// reference's CUDA layer, on the reference's array layouts (one thread per particle).
// Built with -ffp-contract=off: FP64 results follow the reference's operation order.
#include <mutex>

#include "xdem_capi.h"
#include "xdem_contact.hpp"
#include "xdem_runtime.hpp"
#include "xdem_sort.hpp"

namespace xdem {

Runtime& runtime()
{
    static Runtime rt;
    return rt;
}

static std::mutex g_mutex;

namespace {

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Workspace layout for n elements: error word, 1024 reduction partials, 4 sort arrays,
// sort histograms, 6 body-force arrays (each part 256-byte aligned).
struct WsLayout {
    size_t err, red, sort, scratch, bf, total;
};

WsLayout ws_layout(uint32_t n)
{
    WsLayout L;
    size_t o = 0;
    L.err = o; o += 256;
    L.red = o; o += align256(sizeof(double) * 1024);
    L.sort = o; o += align256(sizeof(uint32_t) * 4 * (size_t)n);
    L.scratch = o; o += align256(radix_sort_scratch_bytes(n));
    L.bf = o; o += align256(sizeof(double) * 6 * (size_t)n);
    L.total = o;
    return L;
}

// Points the runtime at workspace `base` for n elements and clears its error word.
hipError_t ws_attach(Runtime& rt, char* base, size_t bytes, uint32_t n, bool owned)
{
    if (rt.ws_owned && rt.ws && rt.ws != base) (void)hipFree(rt.ws);   // replaced: nothing to report
    const WsLayout L = ws_layout(n);
    rt.ws = base;
    rt.ws_bytes = bytes;
    rt.ws_n = n;
    rt.ws_owned = owned;
    rt.err = reinterpret_cast<uint32_t*>(base + L.err);
    rt.red_partial = reinterpret_cast<double*>(base + L.red);
    rt.sort_buf = reinterpret_cast<uint32_t*>(base + L.sort);
    rt.sort_scratch = base + L.scratch;
    rt.own_bf = reinterpret_cast<double*>(base + L.bf);
    hipError_t e = hipMemsetAsync(rt.err, 0, sizeof(uint32_t), rt.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(rt.stream);
    return e;
}

// A library-owned workspace for n elements (an init-time allocation).
hipError_t ws_reserve(Runtime& rt, uint32_t n)
{
    if (rt.ws && (rt.ws_n >= n || !rt.ws_owned)) return hipSuccess;   // enough, or the caller's
    const size_t bytes = ws_layout(n).total;
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return e;
    return ws_attach(rt, static_cast<char*>(p), bytes, n, true);
}

// Status of a launch that needs n elements of workspace.
int ws_require(Runtime& rt, uint32_t n)
{
    if (rt.ws && rt.ws_n >= n) return XDEM_OK;
    rt.host_err |= XDEM_ERR_WORKSPACE;
    return rt.ws ? XDEM_E_INVALID : XDEM_E_NOT_READY;
}

}  // namespace

size_t workspace_bytes(uint32_t n) { return ws_layout(n).total; }

namespace {

constexpr unsigned kThreads = 128;  // the reference's CUDA_THREADS_PER_BLOCK (decl:17)

inline unsigned nblocks(unsigned n) { return (n + kThreads - 1) / kThreads; }

struct LaunchPrm {
    double dt, half2dt;
    double gx, gy, gz;
    Grid grid;
    int model;
};

LaunchPrm launch_prm(const xdem_dem_parameters& p)
{
    LaunchPrm L;
    L.dt = p.dt;
    L.half2dt = p.half2dt;
    L.gx = p.gravity.x; L.gy = p.gravity.y; L.gz = p.gravity.z;
    L.grid.wo_x = p.world_origin.x; L.grid.wo_y = p.world_origin.y; L.grid.wo_z = p.world_origin.z;
    L.grid.cs = p.cell_size;
    L.grid.mask_x = p.grid_size.x - 1; L.grid.mask_y = p.grid_size.y - 1; L.grid.mask_z = p.grid_size.z - 1;
    L.grid.dim_x = p.grid_size.x;
    L.grid.dim_xy = p.grid_size.x * p.grid_size.y;
    L.grid.base = 0;
    L.model = p.contact_model;
    return L;
}

bool pow2(unsigned v) { return v && !(v & (v - 1)); }

// ------------------------------------------------------------ Velocity Verlet (impl:135-222)
__global__ __launch_bounds__(kThreads) void k_vv_position(LaunchPrm L, double* pos, double* ep, const double* vel,
                                                          const double* ev, const double* acc, const double* ea,
                                                          unsigned np)
{
    const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    const V4<double> e = ld4(ep + 4 * id), d = ld4(ev + 4 * id), a = ld4(ea + 4 * id);
    V4<double> ne;
    ne.x = e.x + L.dt * d.x + L.half2dt * a.x;
    ne.y = e.y + L.dt * d.y + L.half2dt * a.y;
    ne.z = e.z + L.dt * d.z + L.half2dt * a.z;
    ne.w = e.w + L.dt * d.w + L.half2dt * a.w;
    const double l = sqrt(ne.x * ne.x + ne.y * ne.y + ne.z * ne.z + ne.w * ne.w);
    ne.x = ne.x / l; ne.y = ne.y / l; ne.z = ne.z / l; ne.w = ne.w / l;
    st4(ep + 4 * id, ne);
    const V3<double> v = ld3(vel + 3 * id), ac = ld3(acc + 3 * id);
    pos[4 * id + 0] += L.dt * v.x + L.half2dt * ac.x;
    pos[4 * id + 1] += L.dt * v.y + L.half2dt * ac.y;
    pos[4 * id + 2] += L.dt * v.z + L.half2dt * ac.z;
}

template <bool KEEP_V>
__global__ __launch_bounds__(kThreads) void k_vv_velocity(LaunchPrm L, double* vel, double* acc, const double* ep,
                                                          double* ev, double* ea, double* force, double* moment,
                                                          const double* mass, const double* iner, unsigned np)
{
    const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    const double m = mass[id];
    const V3<double> F = ld3(force + 3 * id);
    const V4<double> e = ld4(ep + 4 * id);
    V4<double> av = ld4(ev + 4 * id);
    const V4<double> aa = ld4(ea + 4 * id);
    const V3<double> np_ = rot_local(ld3(moment + 3 * id), e);
    const V4<double> mea = uceom(ld3(iner + 3 * id), e, av, np_);
    const double inv_m = 1.0 / m;
    const V3<double> a = mk3(inv_m * (F.x + m * L.gx), inv_m * (F.y + m * L.gy), inv_m * (F.z + m * L.gz));
    const double hdt = 0.5 * L.dt;
    if (KEEP_V) {
        av.x = av.x + hdt * (aa.x + mea.x); av.y = av.y + hdt * (aa.y + mea.y);
        av.z = av.z + hdt * (aa.z + mea.z); av.w = av.w + hdt * (aa.w + mea.w);
    } else {
        V3<double> v = ld3(vel + 3 * id);
        const V3<double> ao = ld3(acc + 3 * id);
        v = mk3(v.x + hdt * ao.x, v.y + hdt * ao.y, v.z + hdt * ao.z);
        av.x = av.x + hdt * aa.x; av.y = av.y + hdt * aa.y; av.z = av.z + hdt * aa.z; av.w = av.w + hdt * aa.w;
        v = mk3(v.x + hdt * a.x, v.y + hdt * a.y, v.z + hdt * a.z);
        av.x = av.x + hdt * mea.x; av.y = av.y + hdt * mea.y; av.z = av.z + hdt * mea.z; av.w = av.w + hdt * mea.w;
        st3(vel + 3 * id, v);
    }
    st4(ev + 4 * id, av);
    st3(acc + 3 * id, a);
    st4(ea + 4 * id, mea);
    st3(force + 3 * id, mk3(0.0, 0.0, 0.0));
    st3(moment + 3 * id, mk3(0.0, 0.0, 0.0));
}

// ------------------------------------------------------------------ hash (impl:290-315)
__global__ __launch_bounds__(kThreads) void k_hash(LaunchPrm L, unsigned* hash, unsigned* index, const double* pos,
                                                   unsigned sid, unsigned np)
{
    const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    int gx, gy, gz;
    grid_cell(L.grid, pos[4 * id], pos[4 * id + 1], pos[4 * id + 2], gx, gy, gz);
    hash[sid + id] = grid_bucket(L.grid, gx, gy, gz);
    index[sid + id] = sid + id;
}

// ------------------------------------------------------------ cell bounds (impl:317-360)
__global__ __launch_bounds__(kThreads) void k_cell_bounds(const unsigned* hash, const unsigned* index, unsigned* cstart,
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

__global__ void k_fill_u32(unsigned* p, unsigned v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// --------------------------------------------------------- sphere-sphere (impl:613-733)
template <int MODEL, bool GPU_LAW, bool STALE>
__global__ __launch_bounds__(kThreads) void k_p2p(LaunchPrm L, const double* __restrict__ pos, const double* __restrict__ ep,
                                                  const double* __restrict__ vel, const double* __restrict__ ev,
                                                  double* force, double* moment, const double* __restrict__ mass,
                                                  double* tmax, double* rres, unsigned* pair_count, unsigned* pair_id,
                                                  double* tsd, const unsigned* __restrict__ sorted_index,
                                                  const unsigned* __restrict__ cstart, const unsigned* __restrict__ cend,
                                                  const xdem_contact_property* cp, unsigned np, unsigned* err)
{
    constexpr unsigned CAP = XDEM_MAX_P2P_COUNT;
    const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    const ContactConst cc = make_contact_const(*cp, 0.0);
    const unsigned sid = id * CAP;
    unsigned pid[CAP];
    double ps0[CAP], ps1[CAP];
#pragma unroll
    for (unsigned s = 0; s < CAP; s++) {
        pid[s] = pair_id[sid + s];
        ps0[s] = tsd[2 * (sid + s)];
        ps1[s] = tsd[2 * (sid + s) + 1];
    }
    unsigned old_count = pair_count[id];
    if (old_count > CAP) old_count = CAP;
    const V4<double> ipos = ld4(pos + 4 * id);
    const V3<double> ivel = ld3(vel + 3 * id);
    const V3<double> iom = omega_of(ld4(ep + 4 * id), ld4(ev + 4 * id));
    int gx, gy, gz;
    grid_cell(L.grid, ipos.x, ipos.y, ipos.z, gx, gy, gz);
    const double ir = ipos.w, im = mass[id];
    V3<double> Ft = mk3(0.0, 0.0, 0.0);
    Accum acc = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), 0.0};
    unsigned n_new = 0, n_seen = 0;
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                const unsigned h = grid_bucket(L.grid, gx + x, gy + y, gz + z);
                const unsigned start = cstart[h];
                if (start == 0xffffffffu) continue;
                const unsigned end = cend[h];
                for (unsigned j = start; j < end; j++) {
                    const unsigned k = sorted_index[j];
                    if (id == k || k >= np) continue;
                    const V4<double> jpos = ld4(pos + 4 * k);
                    const PairGeom g = pair_geom(ipos, jpos);
                    const double jr = jpos.w;
                    double coh_s = 0;
                    if (cc.coh) coh_s = cohesion_limit_depth(cc, jr ? (ir * jr) / (ir + jr) : ir);
                    double sd0 = 0.0, sd1 = 0.0;
                    bool store = false;
                    if (g.cdist > 0) {
                        for (unsigned s = 0; s < old_count; s++)
                            if (pid[s] == k) { sd0 = ps0[s]; sd1 = ps1[s]; break; }
                        const V3<double> jvel = ld3(vel + 3 * k);
                        const V3<double> jom = omega_of(ld4(ep + 4 * k), ld4(ev + 4 * k));
                        if (!STALE) Ft = mk3(0.0, 0.0, 0.0);
                        sphere_contact<double, MODEL, GPU_LAW>(cc, L.dt, ipos, ivel, iom, im, jpos, jvel, jom, mass[k], g,
                                                               sd0, sd1, Ft, acc);
                        store = true;
                    } else if (g.cdist <= 0 && fabs(g.cdist) < fabs(coh_s)) {
                        sphere_cohesion(cc, ir, jr, g, acc);
                        store = true;
                    }
                    if (store) {
                        n_seen++;
                        if (n_new < CAP) {
                            tsd[2 * (sid + n_new)] = sd0;
                            tsd[2 * (sid + n_new) + 1] = sd1;
                            pair_id[sid + n_new] = k;
                            n_new++;
                        }
                    }
                }
            }
    force[3 * id + 0] += acc.F.x; force[3 * id + 1] += acc.F.y; force[3 * id + 2] += acc.F.z;
    moment[3 * id + 0] += acc.M.x; moment[3 * id + 1] += acc.M.y; moment[3 * id + 2] += acc.M.z;
    if (n_seen > CAP) atomicOr(err, XDEM_ERR_P2P_OVERFLOW);
    pair_count[id] = n_new;
    tmax[3 * id + 0] += acc.T.x; tmax[3 * id + 1] += acc.T.y; tmax[3 * id + 2] += acc.T.z;
    rres[id] += acc.R;
}

// ------------------------------------------------------------------ walls (impl:950-1080)
template <int MODEL, bool GPU_LAW>
__global__ __launch_bounds__(kThreads) void k_plane(LaunchPrm L, const xdem_plane_info* plane, const double* dbi,
                                                    const xdem_contact_property* cp, double* fx, double* fy, double* fz,
                                                    double* mx, double* my, double* mz, const double* pos,
                                                    const double* ep, const double* vel, const double* ev,
                                                    double* force, double* moment, const double* mass, double* tmax,
                                                    double* rres, unsigned* pair_count, unsigned* pair_id, double* tsd,
                                                    unsigned np, unsigned* err)
{
    constexpr unsigned CAP = XDEM_MAX_P2PL_COUNT;
    const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    const ContactConst cc = make_contact_const(*cp, 0.0);
    const xdem_plane_info pl = *plane;
    double b[15];
#pragma unroll
    for (int q = 0; q < 15; q++) b[q] = dbi[q];
    const unsigned sid = id * CAP;
    WallSlots ws;
    ws.init(pair_count[id], pair_id + sid, tsd + 2 * sid);
    ws.load();
    const V4<double> ipos = ld4(pos + 4 * id);
    const V3<double> ivel = ld3(vel + 3 * id);
    const V3<double> iom = omega_of(ld4(ep + 4 * id), ld4(ev + 4 * id));
    Accum acc = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), 0.0};
    double bf[6] = {0, 0, 0, 0, 0, 0};
    const bool ovf = wall_contact<double, MODEL, GPU_LAW>(cc, pl, b, L.dt, ipos, ivel, iom, mass[id], ws, acc, bf);
    if (ovf) atomicOr(err, XDEM_ERR_PLANE_OVERFLOW);
#pragma unroll
    for (unsigned s = 0; s < CAP; s++) {
        pair_id[sid + s] = ws.id[s];
        tsd[2 * (sid + s)] = ws.s0[s];
        tsd[2 * (sid + s) + 1] = ws.s1[s];
    }
    pair_count[id] = ws.count;
    fx[id] = bf[0]; fy[id] = bf[1]; fz[id] = bf[2];
    mx[id] = bf[3]; my[id] = bf[4]; mz[id] = bf[5];
    force[3 * id + 0] += acc.F.x; force[3 * id + 1] += acc.F.y; force[3 * id + 2] += acc.F.z;
    moment[3 * id + 0] += acc.M.x; moment[3 * id + 1] += acc.M.y; moment[3 * id + 2] += acc.M.z;
    tmax[3 * id + 0] += acc.T.x; tmax[3 * id + 1] += acc.T.y; tmax[3 * id + 2] += acc.T.z;
    rres[id] += acc.R;
}

// ---------------------------------------------------------------- cylinders (impl:1327-1460)
template <int MODEL, bool GPU_LAW>
__global__ __launch_bounds__(kThreads) void k_cylinder(LaunchPrm L, const xdem_cylinder_info* cyl, const double* dbi,
                                                       const xdem_contact_property* cp, double* fx, double* fy,
                                                       double* fz, double* mx, double* my, double* mz,
                                                       const double* pos, const double* ep, const double* vel,
                                                       const double* ev, double* force, double* moment,
                                                       const double* mass, double* tmax, double* rres,
                                                       unsigned* pair_count, unsigned* pair_id, double* tsd,
                                                       unsigned np, unsigned* err)
{
    constexpr unsigned CAP = XDEM_MAX_P2CY_COUNT;
    const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    const ContactConst cc = make_contact_const(*cp, 0.0);
    const xdem_cylinder_info cy = *cyl;
    double b[15];
#pragma unroll
    for (int q = 0; q < 15; q++) b[q] = dbi[q];
    const unsigned sid = id * CAP;
    CylSlots cs;
#pragma unroll
    for (unsigned s = 0; s < CAP; s++) {
        cs.old_id[s] = cs.id[s] = pair_id[sid + s];
        cs.old_s0[s] = cs.s0[s] = tsd[2 * (sid + s)];
        cs.old_s1[s] = cs.s1[s] = tsd[2 * (sid + s) + 1];
    }
    cs.count = pair_count[id];
    cs.overflow = false;
    const V4<double> ipos = ld4(pos + 4 * id);
    const V3<double> ivel = ld3(vel + 3 * id);
    const V3<double> iom = omega_of(ld4(ep + 4 * id), ld4(ev + 4 * id));
    Accum acc = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), 0.0};
    double bf[6] = {0, 0, 0, 0, 0, 0};
    cylinder_contact<double, MODEL, GPU_LAW>(cc, cy, b, L.dt, ipos, ivel, iom, mass[id], cs, acc, bf);
    if (cs.overflow) atomicOr(err, XDEM_ERR_CYL_OVERFLOW);
#pragma unroll
    for (unsigned s = 0; s < CAP; s++) {
        pair_id[sid + s] = cs.id[s];
        tsd[2 * (sid + s)] = cs.s0[s];
        tsd[2 * (sid + s) + 1] = cs.s1[s];
    }
    pair_count[id] = cs.count;
    fx[id] = bf[0]; fy[id] = bf[1]; fz[id] = bf[2];
    mx[id] = bf[3]; my[id] = bf[4]; mz[id] = bf[5];
    force[3 * id + 0] += acc.F.x; force[3 * id + 1] += acc.F.y; force[3 * id + 2] += acc.F.z;
    moment[3 * id + 0] += acc.M.x; moment[3 * id + 1] += acc.M.y; moment[3 * id + 2] += acc.M.z;
    tmax[3 * id + 0] += acc.T.x; tmax[3 * id + 1] += acc.T.y; tmax[3 * id + 2] += acc.T.z;
    rres[id] += acc.R;
}

// -------------------------------------------------------------- triangles (impl:1930-2109)
constexpr unsigned kTriList = 20;

template <int MODEL, bool GPU_LAW>
__global__ __launch_bounds__(kThreads) void k_mesh(LaunchPrm L, const xdem_triangle_info* dti, const double* dbi,
                                                   double* fx, double* fy, double* fz, double* mx, double* my,
                                                   double* mz, const double* pos, const double* ep, const double* vel,
                                                   const double* ev, double* force, double* moment, const double* mass,
                                                   double* tmax, double* rres, unsigned* pair_count, unsigned* pair_id,
                                                   double* tsd, const unsigned* sorted_index, const unsigned* cstart,
                                                   const unsigned* cend, const xdem_contact_property* cp, unsigned np,
                                                   unsigned ntriangle, unsigned* err)
{
    constexpr unsigned CAP = XDEM_MAX_P2MS_COUNT;
    const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    const ContactConst cc = make_contact_const(*cp, 0.0);
    double b[15];
#pragma unroll
    for (int q = 0; q < 15; q++) b[q] = dbi[q];
    const unsigned sid = id * CAP;
    unsigned pid[CAP];
    double ps0[CAP], ps1[CAP];
#pragma unroll
    for (unsigned s = 0; s < CAP; s++) {
        pid[s] = pair_id[sid + s];
        ps0[s] = tsd[2 * (sid + s)];
        ps1[s] = tsd[2 * (sid + s) + 1];
    }
    const double im = mass[id];
    const V3<double> ipos = ld3(pos + 4 * id);
    const V3<double> ivel = ld3(vel + 3 * id);
    const V3<double> iom = omega_of(ld4(ep + 4 * id), ld4(ev + 4 * id));
    int gx, gy, gz;
    grid_cell(L.grid, ipos.x, ipos.y, ipos.z, gx, gy, gz);
    const double ir = pos[4 * id + 3];
    const unsigned sk = dti[0].tid;
    TriHit ctri[kTriList], cline[kTriList], cpnt[kTriList];
    unsigned nct = 0, ncl = 0, ncp = 0;
    bool list_full = false;
    double coh_s = 0;
    V3<double> prev_line = mk3(0.0, 0.0, 0.0), prev_point = mk3(0.0, 0.0, 0.0);
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                const unsigned h = grid_bucket(L.grid, gx + x, gy + y, gz + z);
                const unsigned start = cstart[h];
                if (start == 0xffffffffu) continue;
                const unsigned end = cend[h];
                for (unsigned j = start; j < end; j++) {
                    unsigned tk = sorted_index[j];
                    if (!(tk >= np + sk && tk < np + sk + ntriangle)) continue;
                    tk -= np + sk;
                    int t;
                    const xdem_triangle_info tri = dti[tk];
                    const V3<double> cpt = closest_on_triangle(tri, ipos, t);
                    const double cdist = ir - length(ipos - cpt);
                    if (cc.coh) coh_s = cohesion_limit_depth(cc, ir);
                    bool take = false;
                    if (cdist > 0) {
                        if (t == 1) {
                            if (same_point(prev_line, cpt)) continue;
                            prev_line = cpt;
                        } else if (t == 0) {
                            if (same_point(prev_point, cpt)) continue;
                            prev_point = cpt;
                        }
                        take = true;
                    } else if (cdist <= 0 && fabs(cdist) < fabs(coh_s)) {
                        take = true;
                    }
                    if (take) {
                        const TriHit hit = {tk, cpt};
                        if (t == 2) { if (nct < kTriList) ctri[nct++] = hit; else list_full = true; }
                        else if (t == 1) { if (ncl < kTriList) cline[ncl++] = hit; else list_full = true; }
                        else if (t == 0) { if (ncp < kTriList) cpnt[ncp++] = hit; else list_full = true; }
                    }
                }
            }
    if (list_full) atomicOr(err, XDEM_ERR_MESH_LIST);
    Accum acc = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), 0.0};
    double bf[6] = {0, 0, 0, 0, 0, 0};
    unsigned n_new = 0, n_seen = 0;
    auto apply = [&](const TriHit& h) {
        double sd0, sd1;
        if (triangle_contact<double, MODEL, GPU_LAW>(cc, b, L.dt, h, ipos, ir, im, ivel, iom, pid, ps0, ps1, sd0, sd1,
                                                     acc, bf)) {
            n_seen++;
            if (n_new < CAP) {
                tsd[2 * (sid + n_new)] = sd0;
                tsd[2 * (sid + n_new) + 1] = sd1;
                pair_id[sid + n_new] = h.id;
                n_new++;
            }
        }
    };
    for (unsigned k = 0; k < nct; k++) apply(ctri[k]);
    if (!nct)
        for (unsigned k = 0; k < ncl; k++) apply(cline[k]);
    if (!nct && !ncl)
        for (unsigned k = 0; k < ncp; k++) apply(cpnt[k]);
    fx[id] = bf[0]; fy[id] = bf[1]; fz[id] = bf[2];
    mx[id] = bf[3]; my[id] = bf[4]; mz[id] = bf[5];
    force[3 * id + 0] += acc.F.x; force[3 * id + 1] += acc.F.y; force[3 * id + 2] += acc.F.z;
    moment[3 * id + 0] += acc.M.x; moment[3 * id + 1] += acc.M.y; moment[3 * id + 2] += acc.M.z;
    if (n_seen > CAP) atomicOr(err, XDEM_ERR_MESH_OVERFLOW);
    pair_count[id] = n_new;
    tmax[3 * id + 0] += acc.T.x; tmax[3 * id + 1] += acc.T.y; tmax[3 * id + 2] += acc.T.z;
    rres[id] += acc.R;
}

// ---------------------------------------------------------------- rolling (impl:2420-2454)
template <bool FLAT>
__global__ __launch_bounds__(kThreads) void k_rolling(LaunchPrm L, double* tmax, double* rres, const double* inertia,
                                                      const double* ep, const double* ev, double* moment, unsigned np)
{
    const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    const double J = FLAT ? inertia[id] : inertia[3 * id];
    V3<double> M = ld3(moment + 3 * id);
    const V3<double> w = omega_of(ld4(ep + 4 * id), ld4(ev + 4 * id));
    if (rolling_moment(L.dt, J, w, ld3(tmax + 3 * id), rres[id], M)) {
        st3(moment + 3 * id, M);
        st3(tmax + 3 * id, mk3(0.0, 0.0, 0.0));
        rres[id] = 0.0;
    }
}

// --------------------------------------------------------------- mesh pose (impl:2544-2601)
__global__ __launch_bounds__(kThreads) void k_mesh_update(const double* dbi, const double* vList, double* sphere,
                                                          const double* dlocal, xdem_triangle_info* dpi, unsigned nt)
{
    const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nt) return;
    const V3<double> bp = mk3(dbi[XDEM_PX], dbi[XDEM_PY], dbi[XDEM_PZ]);
    const V4<double> e = {dbi[XDEM_E0], dbi[XDEM_E1], dbi[XDEM_E2], dbi[XDEM_E3]};
    const double* s = vList + 9 * id;
    const V3<double> P = bp + rot_global(ld3(s + 0), e);
    const V3<double> Q = bp + rot_global(ld3(s + 3), e);
    const V3<double> R = bp + rot_global(ld3(s + 6), e);
    const V3<double> c = bp + rot_global(ld3(dlocal + 3 * id), e);
    sphere[4 * id + 0] = c.x; sphere[4 * id + 1] = c.y; sphere[4 * id + 2] = c.z;
    dpi[id].P = {P.x, P.y, P.z};
    dpi[id].Q = {Q.x, Q.y, Q.z};
    dpi[id].R = {R.x, R.y, R.z};
}

// ------------------------------------------------------------- reduction (def:285-317)
// Fixed-order two-stage FP64 sum: 1024 partial sums of grid-strided elements, then one
// work-group adds the partials in index order.
__global__ __launch_bounds__(256) void k_reduce_partial(const double* in, unsigned n, double* partial)
{
    __shared__ double s[4];
    double v = 0.0;
    for (unsigned i = blockIdx.x * 256 + threadIdx.x; i < n; i += 256 * gridDim.x) v += in[i];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = ((s[0] + s[1]) + s[2]) + s[3];
}

__global__ __launch_bounds__(64) void k_reduce_final(const double* partial, unsigned np, double* out)
{
    double v = 0.0;
    for (unsigned i = threadIdx.x; i < np; i += 64) v += partial[i];
    v = wave_sum(v);
    if (threadIdx.x == 0) out[0] = v;
}

}  // namespace
}  // namespace xdem

using namespace xdem;

#define XDEM_REQUIRE_PARAMS()                           \
    if (!runtime().prm_set) return XDEM_E_NOT_READY;    \
    const LaunchPrm L = launch_prm(runtime().prm)

#define XDEM_LAUNCH_CHECK() \
    do { if (hipGetLastError() != hipSuccess) return XDEM_E_HIP; } while (0)

extern "C" {

int xdem_setDEMSymbolicParameter(const xdem_dem_parameters* h)
{
    if (!h) return XDEM_E_INVALID;
    if (!pow2(h->grid_size.x) || !pow2(h->grid_size.y) || !pow2(h->grid_size.z)) return XDEM_E_INVALID;
    if (!(h->cell_size > 0.0)) return XDEM_E_INVALID;
    std::lock_guard<std::mutex> lk(g_mutex);
    runtime().prm = *h;
    runtime().prm_set = true;
    // the launchers' workspace for the particles and the mesh proxy spheres the caller
    // counts in nsphere, allocated here, at set-up (a caller-owned workspace is kept)
    const unsigned long long n = (unsigned long long)h->np + h->nsphere;
    if (n > 0xffffffffull) return XDEM_E_INVALID;
    XDEM_HIP_CHECK(ws_reserve(runtime(), (uint32_t)n));
    return XDEM_OK;
}

int xdem_set_quirks(const xdem_quirks* q)
{
    if (!q) return XDEM_E_INVALID;
    runtime().quirks = *q;
    return XDEM_OK;
}

int xdem_set_stream(void* s)
{
    runtime().stream = static_cast<hipStream_t>(s);
    return XDEM_OK;
}

int xdem_set_body_force_buffers(double* fx, double* fy, double* fz, double* mx, double* my, double* mz)
{
    double* b[6] = {fx, fy, fz, mx, my, mz};
    for (int k = 0; k < 6; k++) runtime().bf[k] = b[k];
    return XDEM_OK;
}

int xdem_get_body_force_buffers(double** fx, double** fy, double** fz, double** mx, double** my, double** mz)
{
    double** o[6] = {fx, fy, fz, mx, my, mz};
    Runtime& rt = runtime();
    for (int k = 0; k < 6; k++) {
        if (!o[k]) return XDEM_E_INVALID;
        *o[k] = rt.bf[k] ? rt.bf[k] : (rt.own_bf ? rt.own_bf + (size_t)k * rt.ws_n : nullptr);
    }
    return XDEM_OK;
}

int xdem_get_error_word(unsigned int* out, int clear)
{
    if (!out) return XDEM_E_INVALID;
    std::lock_guard<std::mutex> lk(g_mutex);
    Runtime& rt = runtime();
    unsigned dev_word = 0;
    if (rt.err) {
        XDEM_HIP_CHECK(hipMemcpyAsync(&dev_word, rt.err, sizeof(unsigned), hipMemcpyDeviceToHost, rt.stream));
        XDEM_HIP_CHECK(hipStreamSynchronize(rt.stream));
        if (clear) {
            XDEM_HIP_CHECK(hipMemsetAsync(rt.err, 0, sizeof(unsigned), rt.stream));
            XDEM_HIP_CHECK(hipStreamSynchronize(rt.stream));
        }
    }
    *out = dev_word | rt.host_err;
    if (clear) rt.host_err = 0;
    return XDEM_OK;
}

int xdem_workspace_size(unsigned int n, unsigned long long* bytes)
{
    if (!bytes) return XDEM_E_INVALID;
    *bytes = (unsigned long long)workspace_bytes(n);
    return XDEM_OK;
}

int xdem_reserve_workspace(unsigned int n)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    XDEM_HIP_CHECK(ws_reserve(runtime(), n));
    return XDEM_OK;
}

int xdem_set_workspace(void* d_ptr, unsigned long long bytes, unsigned int n)
{
    if (!d_ptr) {
        // detach: the library no longer uses the caller's buffer (the next init call reserves
        // a workspace of its own)
        std::lock_guard<std::mutex> lk(g_mutex);
        Runtime& rt = runtime();
        if (rt.ws_owned && rt.ws) (void)hipFree(rt.ws);   // nothing to report on release
        rt.ws = nullptr; rt.ws_bytes = 0; rt.ws_n = 0; rt.ws_owned = false;
        rt.err = nullptr; rt.red_partial = nullptr; rt.sort_buf = nullptr; rt.sort_scratch = nullptr;
        rt.own_bf = nullptr;
        return XDEM_OK;
    }
    if (reinterpret_cast<uintptr_t>(d_ptr) & 255u) return XDEM_E_INVALID;
    if (bytes < (unsigned long long)workspace_bytes(n)) return XDEM_E_INVALID;
    std::lock_guard<std::mutex> lk(g_mutex);
    XDEM_HIP_CHECK(ws_attach(runtime(), static_cast<char*>(d_ptr), (size_t)bytes, n, false));
    return XDEM_OK;
}

int xdem_vv_update_position(double* pos, double* ep, double* vel, double* ev, double* acc, double* ea, unsigned int np)
{
    XDEM_REQUIRE_PARAMS();
    if (np == 0) return XDEM_OK;
    if (!pos || !ep || !vel || !ev || !acc || !ea) return XDEM_E_INVALID;
    hipLaunchKernelGGL(k_vv_position, dim3(nblocks(np)), dim3(kThreads), 0, runtime().stream, L, pos, ep, vel, ev, acc,
                       ea, np);
    XDEM_LAUNCH_CHECK();
    return XDEM_OK;
}

int xdem_vv_update_velocity(double* vel, double* acc, double* ep, double* ev, double* ea, double* force, double* moment,
                            double* mass, double* iner, unsigned int np)
{
    XDEM_REQUIRE_PARAMS();
    if (np == 0) return XDEM_OK;
    if (!vel || !acc || !ep || !ev || !ea || !force || !moment || !mass || !iner) return XDEM_E_INVALID;
    if (runtime().quirks.velocity_not_updated)
        hipLaunchKernelGGL(k_vv_velocity<true>, dim3(nblocks(np)), dim3(kThreads), 0, runtime().stream, L, vel, acc, ep,
                           ev, ea, force, moment, mass, iner, np);
    else
        hipLaunchKernelGGL(k_vv_velocity<false>, dim3(nblocks(np)), dim3(kThreads), 0, runtime().stream, L, vel, acc, ep,
                           ev, ea, force, moment, mass, iner, np);
    XDEM_LAUNCH_CHECK();
    return XDEM_OK;
}

int xdem_cu_calculateHashAndIndex(unsigned int* hash, unsigned int* index, double* pos, unsigned int sid, unsigned int np)
{
    XDEM_REQUIRE_PARAMS();
    if (np == 0) return XDEM_OK;
    if (!hash || !index || !pos) return XDEM_E_INVALID;
    hipLaunchKernelGGL(k_hash, dim3(nblocks(np)), dim3(kThreads), 0, runtime().stream, L, hash, index, pos, sid, np);
    XDEM_LAUNCH_CHECK();
    return XDEM_OK;
}

int xdem_cu_calculateHashAndIndexForPolygonSphere(unsigned int* hash, unsigned int* index, unsigned int sid,
                                                  unsigned int nsphere, double* sphere)
{
    return xdem_cu_calculateHashAndIndex(hash, index, sphere, sid, nsphere);
}

int xdem_cu_reorderDataAndFindCellStart(unsigned int* hash, unsigned int* index, unsigned int* cstart, unsigned int* cend,
                                        unsigned int* sorted_index, unsigned int np, unsigned int ncell)
{
    if (!hash || !index || !cstart || !cend || !sorted_index) return XDEM_E_INVALID;
    if (ncell == 0) return XDEM_E_INVALID;
    std::lock_guard<std::mutex> lk(g_mutex);
    Runtime& rt = runtime();
    const int rc = ws_require(rt, np);
    if (rc) return rc;
    hipStream_t s = rt.stream;
    if (np > 0) {
        int bits = 1;
        while (bits < 32 && (1ull << bits) < (unsigned long long)ncell) bits++;
        uint32_t* kout = rt.sort_buf;
        uint32_t* vout = rt.sort_buf + rt.ws_n;
        uint32_t* ktmp = rt.sort_buf + 2 * (size_t)rt.ws_n;
        uint32_t* vtmp = rt.sort_buf + 3 * (size_t)rt.ws_n;
        XDEM_HIP_CHECK(radix_sort_pairs(hash, index, kout, vout, ktmp, vtmp, np, bits, rt.sort_scratch, s));
        XDEM_HIP_CHECK(hipMemcpyAsync(hash, kout, sizeof(unsigned) * np, hipMemcpyDeviceToDevice, s));
        XDEM_HIP_CHECK(hipMemcpyAsync(index, vout, sizeof(unsigned) * np, hipMemcpyDeviceToDevice, s));
    }
    // the reference clears both tables every call (def:113-114)
    const unsigned fb = (unsigned)((ncell + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3(fb), dim3(256), 0, s, cstart, 0xffffffffu, (size_t)ncell);
    hipLaunchKernelGGL(k_fill_u32, dim3(fb), dim3(256), 0, s, cend, 0u, (size_t)ncell);
    if (np > 0)
        hipLaunchKernelGGL(k_cell_bounds, dim3(nblocks(np)), dim3(kThreads), 0, s, hash, index, cstart, cend,
                           sorted_index, np);
    XDEM_LAUNCH_CHECK();
    return XDEM_OK;
}

int xdem_cu_calculate_p2p(double* pos, double* ep, double* vel, double* ev, double* force, double* moment, double* mass,
                          double* tmax, double* rres, unsigned int* pair_count, unsigned int* pair_id, double* tsd,
                          unsigned int* sorted_index, unsigned int* cstart, unsigned int* cend, xdem_contact_property* cp,
                          unsigned int np)
{
    XDEM_REQUIRE_PARAMS();
    if (np == 0) return XDEM_OK;
    if (!pos || !ep || !vel || !ev || !force || !moment || !mass || !tmax || !rres || !pair_count || !pair_id || !tsd ||
        !sorted_index || !cstart || !cend || !cp)
        return XDEM_E_INVALID;
    {
        std::lock_guard<std::mutex> lk(g_mutex);
        const int rc = ws_require(runtime(), 0);
        if (rc) return rc;
    }
    const xdem_quirks q = runtime().quirks;
    const dim3 g(nblocks(np)), b(kThreads);
    hipStream_t s = runtime().stream;
    unsigned* err = runtime().err;
#define XDEM_P2P(M, G, S)                                                                                            \
    hipLaunchKernelGGL((k_p2p<M, G, S>), g, b, 0, s, L, pos, ep, vel, ev, force, moment, mass, tmax, rres, pair_count, \
                       pair_id, tsd, sorted_index, cstart, cend, cp, np, err)
    const bool gl = q.gpu_force_law != 0, st = q.stale_ft != 0;
    if (L.model == 1) {
        if (gl) { if (st) XDEM_P2P(1, true, true); else XDEM_P2P(1, true, false); }
        else { if (st) XDEM_P2P(1, false, true); else XDEM_P2P(1, false, false); }
    } else {
        if (gl) { if (st) XDEM_P2P(0, true, true); else XDEM_P2P(0, true, false); }
        else { if (st) XDEM_P2P(0, false, true); else XDEM_P2P(0, false, false); }
    }
#undef XDEM_P2P
    XDEM_LAUNCH_CHECK();
    return XDEM_OK;
}

static int body_force_ptrs(unsigned np, double** bf)
{
    Runtime& rt = runtime();
    bool all = true;
    for (int k = 0; k < 6; k++) all = all && rt.bf[k];
    if (all) {
        for (int k = 0; k < 6; k++) bf[k] = rt.bf[k];
        return XDEM_OK;
    }
    const int rc = ws_require(rt, np);
    if (rc) return rc;
    for (int k = 0; k < 6; k++) bf[k] = rt.own_bf + (size_t)k * rt.ws_n;
    return XDEM_OK;
}

int xdem_cu_plane_contact_force(xdem_plane_info* plan, double* dbi, xdem_contact_property* cp, double* pos, double* ep,
                                double* vel, double* ev, double* force, double* moment, double* mass, double* tmax,
                                double* rres, unsigned int* pair_count, unsigned int* pair_id, double* tsd,
                                unsigned int np)
{
    XDEM_REQUIRE_PARAMS();
    if (np == 0) return XDEM_OK;
    if (!plan || !dbi || !cp || !pos || !ep || !vel || !ev || !force || !moment || !mass || !tmax || !rres ||
        !pair_count || !pair_id || !tsd)
        return XDEM_E_INVALID;
    std::lock_guard<std::mutex> lk(g_mutex);
    int rc0 = ws_require(runtime(), 0);
    if (rc0) return rc0;
    double* bf[6];
    int rc = body_force_ptrs(np, bf);
    if (rc) return rc;
    const bool gl = runtime().quirks.gpu_force_law != 0;
    const dim3 g(nblocks(np)), b(kThreads);
    hipStream_t s = runtime().stream;
    unsigned* err = runtime().err;
#define XDEM_PL(M, G)                                                                                                   \
    hipLaunchKernelGGL((k_plane<M, G>), g, b, 0, s, L, plan, dbi, cp, bf[0], bf[1], bf[2], bf[3], bf[4], bf[5], pos, ep, \
                       vel, ev, force, moment, mass, tmax, rres, pair_count, pair_id, tsd, np, err)
    if (L.model == 1) { if (gl) XDEM_PL(1, true); else XDEM_PL(1, false); }
    else { if (gl) XDEM_PL(0, true); else XDEM_PL(0, false); }
#undef XDEM_PL
    XDEM_LAUNCH_CHECK();
    return XDEM_OK;
}

int xdem_cu_cylinder_contact_force(xdem_cylinder_info* cyl, double* dbi, xdem_contact_property* cp, double* pos,
                                   double* ep, double* vel, double* ev, double* force, double* moment, double* mass,
                                   double* tmax, double* rres, unsigned int* pair_count, unsigned int* pair_id,
                                   double* tsd, unsigned int np)
{
    XDEM_REQUIRE_PARAMS();
    if (np == 0) return XDEM_OK;
    if (!cyl || !dbi || !cp || !pos || !ep || !vel || !ev || !force || !moment || !mass || !tmax || !rres ||
        !pair_count || !pair_id || !tsd)
        return XDEM_E_INVALID;
    std::lock_guard<std::mutex> lk(g_mutex);
    int rc0 = ws_require(runtime(), 0);
    if (rc0) return rc0;
    double* bf[6];
    int rc = body_force_ptrs(np, bf);
    if (rc) return rc;
    const bool gl = runtime().quirks.gpu_force_law != 0;
    const dim3 g(nblocks(np)), b(kThreads);
    hipStream_t s = runtime().stream;
    unsigned* err = runtime().err;
#define XDEM_CY(M, G)                                                                                                  \
    hipLaunchKernelGGL((k_cylinder<M, G>), g, b, 0, s, L, cyl, dbi, cp, bf[0], bf[1], bf[2], bf[3], bf[4], bf[5], pos, \
                       ep, vel, ev, force, moment, mass, tmax, rres, pair_count, pair_id, tsd, np, err)
    if (L.model == 1) { if (gl) XDEM_CY(1, true); else XDEM_CY(1, false); }
    else { if (gl) XDEM_CY(0, true); else XDEM_CY(0, false); }
#undef XDEM_CY
    XDEM_LAUNCH_CHECK();
    return XDEM_OK;
}

int xdem_cu_particle_polygonObject_collision(xdem_triangle_info* dpi, double* dbi, double* pos, double* ep, double* vel,
                                             double* ev, double* force, double* moment, double* mass, double* tmax,
                                             double* rres, unsigned int* pair_count, unsigned int* pair_id, double* tsd,
                                             double* dsph, unsigned int* sidx, unsigned int* cstart, unsigned int* cend,
                                             xdem_contact_property* cp, unsigned int np, unsigned int ntriangle)
{
    XDEM_REQUIRE_PARAMS();
    (void)dsph;
    if (np == 0) return XDEM_OK;
    if (!dpi || !dbi || !pos || !ep || !vel || !ev || !force || !moment || !mass || !tmax || !rres || !pair_count ||
        !pair_id || !tsd || !sidx || !cstart || !cend || !cp || ntriangle == 0)
        return XDEM_E_INVALID;
    std::lock_guard<std::mutex> lk(g_mutex);
    int rc0 = ws_require(runtime(), 0);
    if (rc0) return rc0;
    double* bf[6];
    int rc = body_force_ptrs(np, bf);
    if (rc) return rc;
    const bool gl = runtime().quirks.gpu_force_law != 0;
    const dim3 g(nblocks(np)), b(kThreads);
    hipStream_t s = runtime().stream;
    unsigned* err = runtime().err;
#define XDEM_MS(M, G)                                                                                                  \
    hipLaunchKernelGGL((k_mesh<M, G>), g, b, 0, s, L, dpi, dbi, bf[0], bf[1], bf[2], bf[3], bf[4], bf[5], pos, ep, vel, \
                       ev, force, moment, mass, tmax, rres, pair_count, pair_id, tsd, sidx, cstart, cend, cp, np,     \
                       ntriangle, err)
    if (L.model == 1) { if (gl) XDEM_MS(1, true); else XDEM_MS(1, false); }
    else { if (gl) XDEM_MS(0, true); else XDEM_MS(0, false); }
#undef XDEM_MS
    XDEM_LAUNCH_CHECK();
    return XDEM_OK;
}

int xdem_cu_decide_rolling_friction_moment(double* tmax, double* rres, double* inertia, double* ep, double* ev,
                                           double* moment, unsigned int np)
{
    XDEM_REQUIRE_PARAMS();
    if (np == 0) return XDEM_OK;
    if (!tmax || !rres || !inertia || !ep || !ev || !moment) return XDEM_E_INVALID;
    if (runtime().quirks.flat_inertia)
        hipLaunchKernelGGL(k_rolling<true>, dim3(nblocks(np)), dim3(kThreads), 0, runtime().stream, L, tmax, rres,
                           inertia, ep, ev, moment, np);
    else
        hipLaunchKernelGGL(k_rolling<false>, dim3(nblocks(np)), dim3(kThreads), 0, runtime().stream, L, tmax, rres,
                           inertia, ep, ev, moment, np);
    XDEM_LAUNCH_CHECK();
    return XDEM_OK;
}

int xdem_reduction(double* in, unsigned int np, double* out)
{
    if (!out || (!in && np)) return XDEM_E_INVALID;
    if (np == 0) { *out = 0.0; return XDEM_OK; }
    std::lock_guard<std::mutex> lk(g_mutex);
    Runtime& rt = runtime();
    const int rc = ws_require(rt, 0);
    if (rc) return rc;
    const unsigned nb = std::min(1023u, (np + 255) / 256);
    hipLaunchKernelGGL(k_reduce_partial, dim3(nb), dim3(256), 0, rt.stream, in, np, rt.red_partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, rt.stream, rt.red_partial, nb, rt.red_partial + 1023);
    XDEM_LAUNCH_CHECK();
    XDEM_HIP_CHECK(hipMemcpyAsync(out, rt.red_partial + 1023, sizeof(double), hipMemcpyDeviceToHost, rt.stream));
    XDEM_HIP_CHECK(hipStreamSynchronize(rt.stream));
    return XDEM_OK;
}

int xdem_cu_update_meshObjectData(double* vList, double* sph, double* dlocal, xdem_triangle_info* poly, double* dbi,
                                  unsigned int ntriangle)
{
    if (ntriangle == 0) return XDEM_OK;
    if (!vList || !sph || !dlocal || !poly || !dbi) return XDEM_E_INVALID;
    hipLaunchKernelGGL(k_mesh_update, dim3(nblocks(ntriangle)), dim3(kThreads), 0, runtime().stream, dbi, vList, sph,
                       dlocal, poly, ntriangle);
    XDEM_LAUNCH_CHECK();
    return XDEM_OK;
}

}  // extern "C"

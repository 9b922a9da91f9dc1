// xdem_step.hpp -- kernels of the fused DEM step (the xdem_ctx path).
//
// One step is four kernels plus the sort:
//   k_integrate_hash  Velocity-Verlet position/orientation update (impl:135-147) and the
//                     bucket key of every sphere; clears the buckets the sphere occupied
//                     in the previous step (so no per-step clear of the whole table).
//   radix sort        (key, sphere id) pairs, stable: buckets in key order, spheres of a
//                     bucket in ascending id order, exactly the reference's sorted order.
//   k_cells_gather    bucket ranges [start, end) and a sorted-order copy of what the
//                     neighbours of a sphere read: position+radius, velocity+mass, omega.
//   k_collide         one thread per sorted slot: sphere-sphere contacts over the 27
//                     neighbour buckets in the reference's (z, y, x) order, every wall of
//                     the scene in id order, then (fused mode) rolling resistance and the
//                     velocity/orientation-rate update (xIntegratorVV.cpp:131-152).
// Several independent scenes share the launches: each scene owns a range of buckets, so
// spheres of different scenes never meet in a bucket.
#pragma once

#include "xdem_contact.hpp"

namespace xdem {

constexpr uint32_t kInvalid = 0xffffffffu;
constexpr int kMaxWallsPerScene = 16;

struct SceneDev {
    double dt, half2dt;
    double gx, gy, gz;
    Grid grid;
    ContactConst cc;          // sphere-sphere material pair
    uint32_t p_begin, p_end;  // sphere range of the scene
    uint32_t wall_begin, nwall;
    uint32_t has_pp;
    uint32_t pad_;
};

struct WallDev {
    xdem_plane_info info;
    double dbi[16];           // 15 used
    ContactConst cc;
    int32_t wslot;            // per-sphere wrench slot of a dynamic wall, -1 for a fixed wall
    uint32_t pad_;
};

struct StepArgs {
    // state (reference layouts, all scenes concatenated)
    double* pos;  double* vel;  double* acc;  double* ep;  double* ev;  double* ea;
    double* force; double* moment; const double* mass; const double* inertia;
    double* tmax; double* rres;
    const uint32_t* scene_of;  // sphere -> scene (multi-scene only)
    const SceneDev* scenes;
    const WallDev* walls;
    // grid
    uint32_t* keys; uint32_t* vals;          // unsorted (key, id) of this step
    const uint32_t* skeys; const uint32_t* svals;  // sorted
    uint2* cell_range;                       // [start, end) per bucket, start = kInvalid if empty
    // sorted-order neighbour data
    double4* spos;
    float4* sposf;  // position relative to the scene's grid origin and radius, FP32 (prefilter)
    double4* svm;  // velocity xyz + mass
    double4* som;  // angular velocity xyz
    // sphere history (double-buffered) and wall history (in place)
    const uint32_t* pp_count_in; const uint32_t* pp_id_in; const double2* pp_tsd_in;
    uint32_t* pp_count_out; uint32_t* pp_id_out; double2* pp_tsd_out;
    uint32_t* pl_count; uint32_t* pl_id; double2* pl_tsd;
    double* wrench;          // [nwslot][6][np] per-sphere contributions to dynamic walls
    uint32_t* err;
    unsigned long long* counters;  // [0] stored sphere contacts, [1] stored wall contacts
    uint32_t np, nsorted, cap;
    uint32_t nwslot;
    int stale_ft, flat_inertia, keep_velocity, tmax_persist;
    int zero_force;  // k_finish leaves force/moment zero (the reference's state after a step)
};

__device__ __forceinline__ uint32_t scene_index(const StepArgs& A, uint32_t i, bool multi)
{
    return multi ? A.scene_of[i] : 0u;
}

// ------------------------------------------------------------------- position + key
template <bool MULTI>
__global__ __launch_bounds__(256) void k_integrate_hash(StepArgs A)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {  // stored-contact counters of this step
        A.counters[0] = 0ull;
        A.counters[1] = 0ull;
    }
    if (i >= A.np) return;
    const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
    const uint32_t old = A.keys[i];
    if (old != kInvalid) A.cell_range[old].x = kInvalid;
    const V4<double> e = ld4(A.ep + 4 * i), d = ld4(A.ev + 4 * i), a = ld4(A.ea + 4 * i);
    V4<double> ne;
    ne.x = e.x + S.dt * d.x + S.half2dt * a.x;
    ne.y = e.y + S.dt * d.y + S.half2dt * a.y;
    ne.z = e.z + S.dt * d.z + S.half2dt * a.z;
    ne.w = e.w + S.dt * d.w + S.half2dt * a.w;
    const double l = sqrt(ne.x * ne.x + ne.y * ne.y + ne.z * ne.z + ne.w * ne.w);
    ne.x = ne.x / l; ne.y = ne.y / l; ne.z = ne.z / l; ne.w = ne.w / l;
    st4(A.ep + 4 * i, ne);
    const V3<double> v = ld3(A.vel + 3 * i), ac = ld3(A.acc + 3 * i);
    V4<double> p = ld4(A.pos + 4 * i);
    p.x += S.dt * v.x + S.half2dt * ac.x;
    p.y += S.dt * v.y + S.half2dt * ac.y;
    p.z += S.dt * v.z + S.half2dt * ac.z;
    st4(A.pos + 4 * i, p);
    int gx, gy, gz;
    grid_cell(S.grid, p.x, p.y, p.z, gx, gy, gz);
    A.keys[i] = grid_bucket(S.grid, gx, gy, gz);
    A.vals[i] = i;
}

// --------------------------------------------------------- bucket ranges + sorted copies
// Internal linkage: each translation unit that includes this header gets its own copy;
// only the FP64 unit launches it.
static __global__ __launch_bounds__(256) void k_cells_gather(StepArgs A)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.nsorted) return;
    const uint32_t key = A.skeys[j];
    if (j == 0 || A.skeys[j - 1] != key) A.cell_range[key].x = j;
    if (j == A.nsorted - 1 || A.skeys[j + 1] != key) A.cell_range[key].y = j + 1;
    const uint32_t k = A.svals[j];
    if (k < A.np) {
        const double4 p = reinterpret_cast<const double4*>(A.pos)[k];
        A.spos[j] = p;
        const Grid& gr = A.scenes[A.scene_of ? A.scene_of[k] : 0u].grid;
        A.sposf[j] = make_float4((float)(p.x - gr.wo_x), (float)(p.y - gr.wo_y), (float)(p.z - gr.wo_z), (float)p.w);
        const V3<double> v = ld3(A.vel + 3 * k);
        const V3<double> w = omega_of(ld4(A.ep + 4 * k), ld4(A.ev + 4 * k));
        A.svm[j] = make_double4(v.x, v.y, v.z, A.mass[k]);
        A.som[j] = make_double4(w.x, w.y, w.z, 0.0);
    } else {
        // mesh-proxy slot: infinitely far, never a sphere-sphere candidate
        A.spos[j] = make_double4(__builtin_huge_val(), __builtin_huge_val(), __builtin_huge_val(), 0.0);
        A.sposf[j] = make_float4(__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf(), 0.0f);
    }
}

// ------------------------------------------------------------------------ contacts
template <class T, bool MULTI, bool FUSE, int MODEL, bool GPU_LAW>
__global__ __launch_bounds__(256) void k_collide(StepArgs A)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n_pp = 0, n_wall = 0;
    if (j < A.nsorted) {
        const uint32_t i = A.svals[j];
        if (i < A.np) {
            const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
            const double4 ip4 = A.spos[j];
            const V4<double> ipos = {ip4.x, ip4.y, ip4.z, ip4.w};
            const double4 ivm = A.svm[j];
            const double4 iow = A.som[j];
            const V3<double> ivel = mk3(ivm.x, ivm.y, ivm.z);
            const V3<double> iom = mk3(iow.x, iow.y, iow.z);
            const double im = A.mass[i];
            uint32_t err = 0;

            // ---- sphere-sphere (impl:613-733)
            Accum pp = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), 0.0};
            if (S.has_pp) {
                const uint32_t cap = A.cap;
                const uint32_t hb = i * cap;
                uint32_t old_count = A.pp_count_in[i];
                if (old_count > cap) old_count = cap;
                int gx, gy, gz;
                grid_cell(S.grid, ipos.x, ipos.y, ipos.z, gx, gy, gz);
                V3<T> Ft = mk3(T(0), T(0), T(0));
                uint32_t n_new = 0, n_seen = 0;
                const double reach = S.cc.coh_reach;
                // FP32 prefilter: a candidate is skipped only when its FP32 distance exceeds
                // r_i + r_j + reach by more than the rounding of the FP32 coordinates
                // (|x| 2^-24 per coordinate and operand, bounded by mg) and of the FP32
                // radius sum (factor 1 + 1e-6), so every pair the exact FP64 test below
                // could accept reaches it. Mesh-proxy slots sit at infinity.
                const float4 fi = A.sposf[j];
                const float mg = 2.4e-7f * (fabsf(fi.x) + fabsf(fi.y) + fabsf(fi.z)) + 1e-9f;
                const float reachf = (float)reach;
                // candidates of one run of sorted slots [m0, m1)
                auto visit = [&](uint32_t m0, uint32_t m1) {
                    for (uint32_t m = m0; m < m1; m++) {
                        if (m == j) continue;
                        {
                            const float4 f = A.sposf[m];
                            const float dx = f.x - fi.x, dy = f.y - fi.y, dz = f.z - fi.z;
                            const float lim = (fi.w + f.w + reachf) * 1.000001f + mg;
                            if (dx * dx + dy * dy + dz * dz > lim * lim) continue;
                        }
                        const double4 q = A.spos[m];
                        const uint32_t k = A.svals[m];
                        const V4<double> jpos = {q.x, q.y, q.z, q.w};
                        const PairGeom g = pair_geom(ipos, jpos);
                        double sd0 = 0.0, sd1 = 0.0;
                        bool store = false;
                        if (g.cdist > 0) {
                            for (uint32_t s = 0; s < old_count; s++)
                                if (A.pp_id_in[hb + s] == k) {
                                    const double2 h = A.pp_tsd_in[hb + s];
                                    sd0 = h.x; sd1 = h.y;
                                    break;
                                }
                            const double4 jvm = A.svm[m];
                            const double4 jow = A.som[m];
                            if (!A.stale_ft) Ft = mk3(T(0), T(0), T(0));
                            sphere_contact<T, MODEL, GPU_LAW>(S.cc, S.dt, ipos, ivel, iom, im, jpos,
                                                              mk3(jvm.x, jvm.y, jvm.z), mk3(jow.x, jow.y, jow.z),
                                                              jvm.w, g, sd0, sd1, Ft, pp);
                            store = true;
                        } else if (reach > 0.0 && -g.cdist < reach) {
                            const double jr = jpos.w, ir = ipos.w;
                            const double coh_s = cohesion_limit_depth(S.cc, jr ? (ir * jr) / (ir + jr) : ir);
                            if (fabs(g.cdist) < fabs(coh_s)) {
                                sphere_cohesion(S.cc, ir, jr, g, pp);
                                store = true;
                            }
                        }
                        if (store) {
                            n_seen++;
                            if (n_new < cap) {
                                A.pp_id_out[hb + n_new] = k;
                                A.pp_tsd_out[hb + n_new] = make_double2(sd0, sd1);
                                n_new++;
                            }
                        }
                    }
                };
                // 27 neighbour buckets in the reference's (z, y, x) order. The three buckets
                // of an x-row have consecutive keys unless the row wraps, so their spheres
                // form one run of sorted slots, visited in the same order.
                const bool row_contiguous = (((uint32_t)(gx - 1) & S.grid.mask_x) + 2u) == ((uint32_t)(gx + 1) & S.grid.mask_x);
                for (int z = -1; z <= 1; z++)
                    for (int y = -1; y <= 1; y++) {
                        const uint32_t b0 = grid_bucket(S.grid, gx - 1, gy + y, gz + z);
                        if (row_contiguous) {
                            const uint2 r0 = A.cell_range[b0], r1 = A.cell_range[b0 + 1], r2 = A.cell_range[b0 + 2];
                            const uint32_t s = r0.x != kInvalid ? r0.x : (r1.x != kInvalid ? r1.x : r2.x);
                            if (s == kInvalid) continue;
                            const uint32_t e = r2.x != kInvalid ? r2.y : (r1.x != kInvalid ? r1.y : r0.y);
                            visit(s, e);
                        } else {
                            for (int x = -1; x <= 1; x++) {
                                const uint2 rg = A.cell_range[grid_bucket(S.grid, gx + x, gy + y, gz + z)];
                                if (rg.x != kInvalid) visit(rg.x, rg.y);
                            }
                        }
                    }
                A.pp_count_out[i] = n_new;
                if (n_seen > cap) err |= XDEM_ERR_P2P_OVERFLOW;
                n_pp = n_new;
            }

            // ---- walls, in id order (impl:950-1080 once per wall)
            Accum tot;
            tot.F = mk3(0.0, 0.0, 0.0) + pp.F;
            tot.M = mk3(0.0, 0.0, 0.0) + pp.M;
            const V3<double> t0 = A.tmax_persist ? ld3(A.tmax + 3 * i) : mk3(0.0, 0.0, 0.0);
            tot.T = t0 + pp.T;
            tot.R = 0.0 + pp.R;
            if (S.nwall) {
                WallSlots ws;
#pragma unroll
                for (int s = 0; s < XDEM_MAX_P2PL_COUNT; s++) {
                    ws.id[s] = A.pl_id[XDEM_MAX_P2PL_COUNT * i + s];
                    const double2 h = A.pl_tsd[XDEM_MAX_P2PL_COUNT * i + s];
                    ws.s0[s] = h.x; ws.s1[s] = h.y;
                }
                ws.count = A.pl_count[i];
                for (uint32_t w = 0; w < S.nwall; w++) {
                    const WallDev& W = A.walls[S.wall_begin + w];
                    double bf[6] = {0, 0, 0, 0, 0, 0};
                    if (wall_contact<T, MODEL, GPU_LAW>(W.cc, W.info, W.dbi, S.dt, ipos, ivel, iom, im, ws, tot, bf))
                        err |= XDEM_ERR_PLANE_OVERFLOW;
                    if (W.wslot >= 0) {
                        double* wr = A.wrench + (size_t)W.wslot * 6 * A.np;
#pragma unroll
                        for (int c = 0; c < 6; c++) wr[(size_t)c * A.np + i] = bf[c];
                    }
                }
#pragma unroll
                for (int s = 0; s < XDEM_MAX_P2PL_COUNT; s++) {
                    A.pl_id[XDEM_MAX_P2PL_COUNT * i + s] = ws.id[s];
                    A.pl_tsd[XDEM_MAX_P2PL_COUNT * i + s] = make_double2(ws.s0[s], ws.s1[s]);
                }
                A.pl_count[i] = ws.count;
                n_wall = ws.count;
            }

            if (FUSE) {
                // ---- rolling resistance (impl:2420-2454)
                const V4<double> e = ld4(A.ep + 4 * i);
                V4<double> av = ld4(A.ev + 4 * i);
                // quirk Q5: the scene's inertia array read as flat doubles at the scene-local id
                const double J = A.flat_inertia ? A.inertia[i + 2 * S.p_begin] : A.inertia[3 * i];
                const bool consumed = rolling_moment(S.dt, J, omega_of(e, av), tot.T, tot.R, tot.M);
                if (A.tmax_persist) st3(A.tmax + 3 * i, consumed ? mk3(0.0, 0.0, 0.0) : tot.T);
                // ---- velocity and orientation-rate update (xIntegratorVV.cpp:131-152; impl:192-222)
                const double m = im;
                const V4<double> aa = ld4(A.ea + 4 * i);
                const V3<double> np_ = rot_local(tot.M, e);
                const V4<double> mea = uceom(ld3(A.inertia + 3 * i), e, av, np_);
                const double inv_m = 1.0 / m;
                const V3<double> a = mk3(inv_m * (tot.F.x + m * S.gx), inv_m * (tot.F.y + m * S.gy),
                                         inv_m * (tot.F.z + m * S.gz));
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
                    av.x = av.x + hdt * mea.x; av.y = av.y + hdt * mea.y;
                    av.z = av.z + hdt * mea.z; av.w = av.w + hdt * mea.w;
                    st3(A.vel + 3 * i, v);
                }
                st4(A.ev + 4 * i, av);
                st3(A.acc + 3 * i, a);
                st4(A.ea + 4 * i, mea);
                if (!(isfinite(a.x) && isfinite(a.y) && isfinite(a.z))) err |= XDEM_ERR_NONFINITE;
            } else {
                st3(A.force + 3 * i, tot.F);
                st3(A.moment + 3 * i, tot.M);
                st3(A.tmax + 3 * i, tot.T);
                A.rres[i] = tot.R;
            }
            if (err) atomicOr(A.err, err);
        }
    }
    // stored-contact counters: one atomic per wave
    unsigned long long c0 = n_pp, c1 = n_wall;
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

// Second half of a split step: rolling resistance and the velocity update from the
// force/moment/tmax/rres arrays (xContactManager.cpp:471, xIntegratorVV.cpp:62).
template <bool MULTI>
__global__ __launch_bounds__(256) void k_finish(StepArgs A)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.np) return;
    const SceneDev& S = A.scenes[scene_index(A, i, MULTI)];
    const V4<double> e = ld4(A.ep + 4 * i);
    V4<double> av = ld4(A.ev + 4 * i);
    V3<double> M = ld3(A.moment + 3 * i);
    const double J = A.flat_inertia ? A.inertia[i + 2 * S.p_begin] : A.inertia[3 * i];
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
}

// Kernel entry points of one precision (defined in xdem_step_f64.hip / xdem_step_mixed.hip).
// Contact kernel of one precision (xdem_step_f64.hip / xdem_step_mixed.hip); the other
// kernels are FP64 only and live in xdem_step_f64.hip.
using CollideLaunch = hipError_t (*)(const StepArgs&, bool multi, bool fuse, int model, bool gpu_law, hipStream_t);
CollideLaunch collide_launch_f64();
CollideLaunch collide_launch_mixed();
hipError_t launch_integrate_hash(const StepArgs&, bool multi, hipStream_t);
hipError_t launch_gather(const StepArgs&, hipStream_t);
hipError_t launch_finish(const StepArgs&, bool multi, hipStream_t);

}  // namespace xdem

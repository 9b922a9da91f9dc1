// xdem_calib.hip -- FP64 forward runs of one scene with checkpoints, and their discrete
// adjoint with respect to the contact material (friction, restitution, modulus scale).
//
// Forward step (one sphere per thread, the reference's operations, xdem_calib.hpp):
//   k_cal_position  position and Euler-parameter update, cell hash (impl:54-72, 135-190)
//   cell sort       xdem_cu_reorderDataAndFindCellStart (the launcher layer's stable sort)
//   k_cal_contacts  sphere-sphere contacts over the 27 cells in the reference's order, then
//                   the walls in id order, rolling resistance and the velocity update; velocity
//                   and rate go to the other buffer of a pair (neighbours read the old ones);
//                   optionally records the step's contacts (the tape) for the reverse sweep.
// Reverse step (the same code on Dual<N>, each kernel's vector-Jacobian product from its local
// Jacobian):
//   k_cal_rev_tail      rolling resistance + velocity update of each sphere
//   k_cal_rev_contacts  every recorded contact of each sphere: cotangents of both spheres'
//                       positions, velocities and angular velocities (atomics for the
//                       neighbour), of the history entry it read, and of theta
//   k_cal_rev_head      angular velocity from (e, e_dot), then the position update
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "xdem_calib.h"
#include "xdem_calib.hpp"
#include "xdem_capi.h"

using namespace xdem;
using namespace xdem::calib;

namespace {

constexpr uint32_t kPlaneTag = 0x80000000u;
constexpr int kTape = XDEM_CALIB_TAPE;
constexpr int kCapMax = XDEM_CALIB_MAX_CAPACITY;
constexpr int kPlSlots = XDEM_MAX_P2PL_COUNT;
constexpr int kThreads = 128;

// One state of the run: what a step reads and writes, plus the step's recorded extras.
struct StateView {
    double *pos, *vel, *acc, *ep, *ev, *ea, *tmax, *rres;
    uint32_t *pp_count, *pp_id;
    double* pp_tsd;
    uint32_t *pl_count, *pl_id;
    double* pl_tsd;
};

struct CalArgs {
    uint32_t n, cap, nplane, ncell;
    double dt, h, gx, gy, gz;
    double wo[3], cs;
    uint32_t gsize[3];
    Law law;
    int flat_inertia;
    Material mat_pp;
    const Material* mat_pl;
    const xdem_plane_info* planes;
    const double* plane_dbi;
    double mu, rest, s;              // theta
    const double* mass;
    const double* inertia;
    StateView st;                    // current state (contact kernel: positions already updated)
    double* vel_out;
    double* ev_out;
    uint32_t *hash, *index;
    const uint32_t *sorted, *cstart, *cend;
    uint32_t* tape_code;             // [n * kTape] or null
    uint16_t* tape_slot;
    uint8_t* tape_n;
    uint32_t* err;
};

// ------------------------------------------------------------------------- small helpers
__host__ __device__ inline V3<double> ld3d(const double* p) { return {p[0], p[1], p[2]}; }
__host__ __device__ inline V4<double> ld4d(const double* p) { return {p[0], p[1], p[2], p[3]}; }

__device__ inline void grid_of(const CalArgs& A, V3<double> x, int g[3])
{
    g[0] = (int)floor((x.x - A.wo[0]) / A.cs);
    g[1] = (int)floor((x.y - A.wo[1]) / A.cs);
    g[2] = (int)floor((x.z - A.wo[2]) / A.cs);
}

__device__ inline uint32_t hash_of(const CalArgs& A, int gx, int gy, int gz)
{
    const uint32_t x = (uint32_t)gx & (A.gsize[0] - 1);
    const uint32_t y = (uint32_t)gy & (A.gsize[1] - 1);
    const uint32_t z = (uint32_t)gz & (A.gsize[2] - 1);
    return (z * A.gsize[1]) * A.gsize[0] + y * A.gsize[0] + x;
}

template <int N> __host__ __device__ inline V3<Dual<N>> seed3(V3<double> v, int k0)
{
    return {Dual<N>::input(v.x, k0), Dual<N>::input(v.y, k0 + 1), Dual<N>::input(v.z, k0 + 2)};
}
template <int N> __host__ __device__ inline V4<Dual<N>> seed4(V4<double> v, int k0)
{
    return {Dual<N>::input(v.x, k0), Dual<N>::input(v.y, k0 + 1), Dual<N>::input(v.z, k0 + 2),
            Dual<N>::input(v.w, k0 + 3)};
}
template <int N> __host__ __device__ inline V3<Dual<N>> const3(V3<double> v) { return {Dual<N>(v.x), Dual<N>(v.y), Dual<N>(v.z)}; }

// cotangent accumulation: out[k] += bar * d/dk of the dual y
template <int N> __host__ __device__ inline void vjp_add(double (&out)[N], const Dual<N>& y, double bar)
{
    if (bar == 0.0) return;
    for (int k = 0; k < N; k++) out[k] += bar * y.d[k];
}
template <int N> __host__ __device__ inline void vjp_add3(double (&out)[N], const V3<Dual<N>>& y, V3<double> bar)
{
    vjp_add(out, y.x, bar.x);
    vjp_add(out, y.y, bar.y);
    vjp_add(out, y.z, bar.z);
}

__device__ inline void atomic_add3(double* p, double a, double b, double c)
{
    if (a != 0.0) atomicAdd(p, a);
    if (b != 0.0) atomicAdd(p + 1, b);
    if (c != 0.0) atomicAdd(p + 2, c);
}

// ----------------------------------------------------------------------------- forward
__global__ __launch_bounds__(kThreads) void k_cal_position(CalArgs A)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const StateView& s = A.st;
    const V4<double> p4 = ld4d(s.pos + 4 * (size_t)i);
    V3<double> p = {p4.x, p4.y, p4.z};
    V4<double> e = ld4d(s.ep + 4 * (size_t)i);
    position_update<double>(A.dt, A.h, p, e, ld3d(s.vel + 3 * (size_t)i), ld4d(s.ev + 4 * (size_t)i),
                            ld3d(s.acc + 3 * (size_t)i), ld4d(s.ea + 4 * (size_t)i));
    s.pos[4 * (size_t)i] = p.x;
    s.pos[4 * (size_t)i + 1] = p.y;
    s.pos[4 * (size_t)i + 2] = p.z;
    s.ep[4 * (size_t)i] = e.x;
    s.ep[4 * (size_t)i + 1] = e.y;
    s.ep[4 * (size_t)i + 2] = e.z;
    s.ep[4 * (size_t)i + 3] = e.w;
    int g[3];
    grid_of(A, p, g);
    A.hash[i] = hash_of(A, g[0], g[1], g[2]);
    A.index[i] = i;
}

__device__ inline void tape_put(const CalArgs& A, uint32_t i, uint32_t& nt, uint32_t code, int slot_in, int slot_out,
                                uint32_t& err)
{
    if (!A.tape_code) return;
    if (nt >= (uint32_t)kTape) {
        err |= XDEM_ERR_P2P_OVERFLOW;
        return;
    }
    A.tape_code[(size_t)i * kTape + nt] = code;
    A.tape_slot[(size_t)i * kTape + nt] = (uint16_t)((slot_in + 1) | ((slot_out + 1) << 8));
    nt++;
}

__global__ __launch_bounds__(kThreads) void k_cal_contacts(CalArgs A)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const StateView& s = A.st;
    const Theta<double> th = {A.mu, A.rest, A.s};
    const V4<double> ip = ld4d(s.pos + 4 * (size_t)i);
    const V3<double> xi = {ip.x, ip.y, ip.z};
    const double ir = ip.w, im = A.mass[i];
    const V3<double> vi = ld3d(s.vel + 3 * (size_t)i);
    const V4<double> ei = ld4d(s.ep + 4 * (size_t)i), di = ld4d(s.ev + 4 * (size_t)i);
    const V3<double> wi = ang_vel(ei, di);
    uint32_t err = 0, nt = 0;

    // ---- sphere-sphere (impl:613-733)
    const uint32_t cap = A.cap;
    const size_t sb = (size_t)i * cap;
    const uint32_t old_count = min(s.pp_count[i], cap);
    uint32_t p_id[kCapMax];
    double p_sd[2 * kCapMax];
    for (uint32_t q = 0; q < cap; q++) {
        p_id[q] = s.pp_id[sb + q];
        p_sd[2 * q] = s.pp_tsd[2 * (sb + q)];
        p_sd[2 * q + 1] = s.pp_tsd[2 * (sb + q) + 1];
    }
    Sums<double> pp = zero_sums<double>();
    V3<double> Ft = {0.0, 0.0, 0.0};
    uint32_t n_new = 0, n_seen = 0;
    int g[3];
    grid_of(A, xi, g);
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                const uint32_t hsh = hash_of(A, g[0] + x, g[1] + y, g[2] + z);
                const uint32_t start = A.cstart[hsh];
                if (start == 0xffffffffu) continue;
                const uint32_t end = A.cend[hsh];
                for (uint32_t jj = start; jj < end; jj++) {
                    const uint32_t k = A.sorted[jj];
                    if (k == i || k >= A.n) continue;
                    const V4<double> jp = ld4d(s.pos + 4 * (size_t)k);
                    const V3<double> xj = {jp.x, jp.y, jp.z};
                    const V3<double> rp = xj - xi;
                    const double cdist = (ir + jp.w) - length(rp);
                    if (!(cdist > 0)) continue;
                    double sd0 = 0.0, sd1 = 0.0;
                    int slot_in = -1;
                    for (uint32_t q = 0; q < old_count; q++)
                        if (p_id[q] == k) {
                            sd0 = p_sd[2 * q];
                            sd1 = p_sd[2 * q + 1];
                            slot_in = (int)q;
                            break;
                        }
                    const V3<double> vj = ld3d(s.vel + 3 * (size_t)k);
                    const V3<double> wj = ang_vel(ld4d(s.ep + 4 * (size_t)k), ld4d(s.ev + 4 * (size_t)k));
                    const PairK<double> c = pair_k<double>(A.law.model, A.mat_pp, th, ir, jp.w, im, A.mass[k]);
                    pair_side<double>(A.law, c, A.mat_pp.rfric, xi, ir, xj, jp.w, vi, vj, wi, wj, sd0, sd1, Ft, pp);
                    n_seen++;
                    int slot_out = -1;
                    if (n_new < cap) {
                        s.pp_tsd[2 * (sb + n_new)] = sd0;
                        s.pp_tsd[2 * (sb + n_new) + 1] = sd1;
                        s.pp_id[sb + n_new] = k;
                        slot_out = (int)n_new;
                        n_new++;
                    }
                    tape_put(A, i, nt, k, slot_in, slot_out, err);
                }
            }
    if (n_seen > cap) err |= XDEM_ERR_P2P_OVERFLOW;
    s.pp_count[i] = n_new;
    V3<double> F = V3<double>{0.0, 0.0, 0.0} + pp.F;
    V3<double> M = V3<double>{0.0, 0.0, 0.0} + pp.M;
    V3<double> Tm = ld3d(s.tmax + 3 * (size_t)i) + pp.T;
    double R = s.rres[i] + pp.R;

    // ---- walls in id order (impl:950-1080, one launch per wall)
    for (uint32_t w = 0; w < A.nplane; w++) {
        const xdem_plane_info& pl = A.planes[w];
        const size_t lb = (size_t)i * kPlSlots;
        uint32_t l_id[kPlSlots];
        double l_sd[2 * kPlSlots];
        for (int q = 0; q < kPlSlots; q++) {
            l_id[q] = s.pl_id[lb + q];
            l_sd[2 * q] = s.pl_tsd[2 * (lb + q)];
            l_sd[2 * q + 1] = s.pl_tsd[2 * (lb + q) + 1];
        }
        const uint32_t kid = pl.id;
        uint32_t nn = kid ? s.pl_count[i] : 0u;
        if (nn > (uint32_t)kPlSlots) nn = kPlSlots;
        Sums<double> ps = zero_sums<double>();
        const V3<double> dp = xi - V3<double>{pl.xw.x, pl.xw.y, pl.xw.z};
        const V3<double> wp = {dot(dp, V3<double>{pl.u1.x, pl.u1.y, pl.u1.z}), dot(dp, V3<double>{pl.u2.x, pl.u2.y, pl.u2.z}),
                               dot(dp, V3<double>{pl.uw.x, pl.uw.y, pl.uw.z})};
        V3<double> u0 = {0.0, 0.0, 0.0};
        const double cd = plane_detect<double>(pl, xi, wp, u0, ir);
        if (cd > 0) {
            double sd0 = 0.0, sd1 = 0.0;
            int slot_in = -1;
            for (int q = 0; q < kPlSlots; q++)
                if (l_id[q] == kid) {
                    sd0 = l_sd[2 * q];
                    sd1 = l_sd[2 * q + 1];
                    slot_in = q;
                    break;
                }
            const PairK<double> c = pair_k<double>(A.law.model, A.mat_pl[w], th, ir, 0.0, im, 0.0);
            plane_side<double>(A.law, c, A.mat_pl[w].rfric, pl, A.plane_dbi + 15 * (size_t)w, xi, ir, vi, wi, sd0, sd1,
                               ps);
            int slot_out = -1;
            if (nn < (uint32_t)kPlSlots) {
                s.pl_tsd[2 * (lb + nn)] = sd0;
                s.pl_tsd[2 * (lb + nn) + 1] = sd1;
                s.pl_id[lb + nn] = kid;
                slot_out = (int)nn;
                nn++;
            } else {
                err |= XDEM_ERR_PLANE_OVERFLOW;
            }
            tape_put(A, i, nt, kPlaneTag | w, slot_in, slot_out, err);
        }
        s.pl_count[i] = nn;
        F = F + ps.F;
        M = M + ps.M;
        Tm = Tm + ps.T;
        R = R + ps.R;
    }
    if (A.tape_n) A.tape_n[i] = (uint8_t)nt;

    // ---- rolling resistance (impl:2420-2454), then the velocity update
    if (R != 0.0) {
        const double J = A.flat_inertia ? A.inertia[i] : A.inertia[3 * (size_t)i];
        rolling<double>(A.dt, J, wi, Tm, R, M);
        Tm = {0.0, 0.0, 0.0};
        R = 0.0;
    }
    s.tmax[3 * (size_t)i] = Tm.x;
    s.tmax[3 * (size_t)i + 1] = Tm.y;
    s.tmax[3 * (size_t)i + 2] = Tm.z;
    s.rres[i] = R;
    V3<double> v = vi, acc = ld3d(s.acc + 3 * (size_t)i);
    V4<double> ev = di, ea = ld4d(s.ea + 4 * (size_t)i);
    velocity_update<double>(A.dt, V3<double>{A.gx, A.gy, A.gz}, im, ld3d(A.inertia + 3 * (size_t)i), ei, F, M, v, acc,
                            ev, ea);
    A.vel_out[3 * (size_t)i] = v.x;
    A.vel_out[3 * (size_t)i + 1] = v.y;
    A.vel_out[3 * (size_t)i + 2] = v.z;
    A.ev_out[4 * (size_t)i] = ev.x;
    A.ev_out[4 * (size_t)i + 1] = ev.y;
    A.ev_out[4 * (size_t)i + 2] = ev.z;
    A.ev_out[4 * (size_t)i + 3] = ev.w;
    s.acc[3 * (size_t)i] = acc.x;
    s.acc[3 * (size_t)i + 1] = acc.y;
    s.acc[3 * (size_t)i + 2] = acc.z;
    s.ea[4 * (size_t)i] = ea.x;
    s.ea[4 * (size_t)i + 1] = ea.y;
    s.ea[4 * (size_t)i + 2] = ea.z;
    s.ea[4 * (size_t)i + 3] = ea.w;
    if (!(isfinite(acc.x) && isfinite(acc.y) && isfinite(acc.z))) err |= XDEM_ERR_NONFINITE;
    if (err) atomicOr(A.err, err);
}

// ----------------------------------------------------------------------------- reverse
// The recorded step k: the state the contact kernel read (positions and Euler parameters after
// the position update; velocities, rates, accelerations as before the step), the state before
// the position update's Euler parameters, the slots before the step and the tape.
struct StepRec {
    const double *pos1, *ep0, *ep1, *vel, *ev, *acc, *ea, *tmax, *rres;
    const uint32_t *pp_count, *pp_id;
    const double* pp_tsd;
    const uint32_t *pl_count, *pl_id;
    const double* pl_tsd;
    const uint32_t* tape_code;
    const uint16_t* tape_slot;
    const uint8_t* tape_n;
};

// Cotangents of a state (the reverse sweep's lambda): positions 3 per sphere.
struct Lam {
    double *pos, *vel, *acc, *ep, *ev, *ea, *pp_tsd, *pl_tsd;
};

struct RevArgs {
    CalArgs A;          // sizes, law, materials, theta, mass, inertia
    StepRec r;
    Lam la;             // after the step (read)
    Lam lb;             // before the step (written)
    double* le1;        // [4n] cotangent of e' from the tail and the angular velocity
    double* lom;        // [3n] cotangent of the angular velocity w(e', e_dot)
    double* lpos1;      // [3n] cotangent of the updated positions from the contacts
    double* bar;        // [10n] cotangents of F (3), contact moment (3), Tm (3), Tr (1)
    double* gtheta;     // [3]
};

// The contact sums of sphere i at step r (the forward's order and values).
struct TailIn {
    V3<double> F, Mc, Tm;
    double R;
};

__device__ inline TailIn recompute_sums(const RevArgs& R, uint32_t i)
{
    const CalArgs& A = R.A;
    const StepRec& r = R.r;
    const Theta<double> th = {A.mu, A.rest, A.s};
    const V4<double> ip = ld4d(r.pos1 + 4 * (size_t)i);
    const V3<double> xi = {ip.x, ip.y, ip.z};
    const double ir = ip.w, im = A.mass[i];
    const V3<double> vi = ld3d(r.vel + 3 * (size_t)i);
    const V3<double> wi = ang_vel(ld4d(r.ep1 + 4 * (size_t)i), ld4d(r.ev + 4 * (size_t)i));
    Sums<double> pp = zero_sums<double>();
    V3<double> Ft = {0.0, 0.0, 0.0};
    const uint32_t nt = r.tape_n[i];
    const size_t sb = (size_t)i * A.cap;
    // sphere entries come first, in the forward's order
    uint32_t m = 0;
    for (; m < nt; m++) {
        const uint32_t code = r.tape_code[(size_t)i * kTape + m];
        if (code & kPlaneTag) break;
        const int slot_in = (int)(r.tape_slot[(size_t)i * kTape + m] & 0xffu) - 1;
        double sd0 = 0.0, sd1 = 0.0;
        if (slot_in >= 0) {
            sd0 = r.pp_tsd[2 * (sb + slot_in)];
            sd1 = r.pp_tsd[2 * (sb + slot_in) + 1];
        }
        const V4<double> jp = ld4d(r.pos1 + 4 * (size_t)code);
        const PairK<double> c = pair_k<double>(A.law.model, A.mat_pp, th, ir, jp.w, im, A.mass[code]);
        pair_side<double>(A.law, c, A.mat_pp.rfric, xi, ir, V3<double>{jp.x, jp.y, jp.z}, jp.w, vi,
                          ld3d(r.vel + 3 * (size_t)code), wi,
                          ang_vel(ld4d(r.ep1 + 4 * (size_t)code), ld4d(r.ev + 4 * (size_t)code)), sd0, sd1, Ft, pp);
    }
    TailIn t;
    t.F = V3<double>{0.0, 0.0, 0.0} + pp.F;
    t.Mc = V3<double>{0.0, 0.0, 0.0} + pp.M;
    t.Tm = ld3d(r.tmax + 3 * (size_t)i) + pp.T;
    t.R = r.rres[i] + pp.R;
    // walls: every wall adds its sums (zero without a contact), in id order
    const size_t lb = (size_t)i * kPlSlots;
    for (uint32_t w = 0; w < A.nplane; w++) {
        Sums<double> ps = zero_sums<double>();
        if (m < nt && r.tape_code[(size_t)i * kTape + m] == (kPlaneTag | w)) {
            const int slot_in = (int)(r.tape_slot[(size_t)i * kTape + m] & 0xffu) - 1;
            double sd0 = 0.0, sd1 = 0.0;
            if (slot_in >= 0) {
                sd0 = r.pl_tsd[2 * (lb + slot_in)];
                sd1 = r.pl_tsd[2 * (lb + slot_in) + 1];
            }
            const PairK<double> c = pair_k<double>(A.law.model, A.mat_pl[w], th, ir, 0.0, im, 0.0);
            plane_side<double>(A.law, c, A.mat_pl[w].rfric, A.planes[w], A.plane_dbi + 15 * (size_t)w, xi, ir, vi, wi,
                               sd0, sd1, ps);
            m++;
        }
        t.F = t.F + ps.F;
        t.Mc = t.Mc + ps.M;
        t.Tm = t.Tm + ps.T;
        t.R = t.R + ps.R;
    }
    return t;
}

// Rolling resistance and the velocity update of sphere i. The translational part and the
// rate update are linear and taken by hand; the Euler parameter accelerations (rolling
// resistance, A(e)^T M, calculate_uceom) are differentiated with respect to e' (0-3), e_dot
// (4-7), the contact moment (8-10), Tm (11-13) and Tr (14).
constexpr int kTailN = 15;

__global__ __launch_bounds__(64) void k_cal_rev_tail(RevArgs R)
{
    const CalArgs& A = R.A;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const StepRec& r = R.r;
    const Lam &la = R.la, &lb = R.lb;
    const TailIn t = recompute_sums(R, i);
    const double hdt = 0.5 * A.dt, inv_m = 1.0 / A.mass[i];
    // v' = (v + hdt acc) + hdt a, acc' = a = (F + m g) / m; ev' = (ev + hdt ea) + hdt mea, ea' = mea
    const V3<double> lv = ld3d(la.vel + 3 * (size_t)i), lacc = ld3d(la.acc + 3 * (size_t)i);
    const V4<double> lev = ld4d(la.ev + 4 * (size_t)i), lea = ld4d(la.ea + 4 * (size_t)i);
    const double lmea[4] = {hdt * lev.x + lea.x, hdt * lev.y + lea.y, hdt * lev.z + lea.z, hdt * lev.w + lea.w};
    typedef Dual<kTailN> D;
    const V4<D> e1 = seed4<kTailN>(ld4d(r.ep1 + 4 * (size_t)i), 0);
    const V4<D> ev = seed4<kTailN>(ld4d(r.ev + 4 * (size_t)i), 4);
    V3<D> M = seed3<kTailN>(t.Mc, 8);
    const V3<D> Tm = seed3<kTailN>(t.Tm, 11);
    const D Tr = D::input(t.R, 14);
    if (t.R != 0.0) {
        const double J = A.flat_inertia ? A.inertia[i] : A.inertia[3 * (size_t)i];
        rolling<D>(A.dt, J, ang_vel(e1, ev), Tm, Tr, M);
    }
    const V4<D> mea = uceom<D>(ld3d(A.inertia + 3 * (size_t)i), e1, ev, to_local(M, e1));
    double g[kTailN];
    for (int k = 0; k < kTailN; k++) g[k] = 0.0;
    vjp_add(g, mea.x, lmea[0]);
    vjp_add(g, mea.y, lmea[1]);
    vjp_add(g, mea.z, lmea[2]);
    vjp_add(g, mea.w, lmea[3]);
    const double lev_a[4] = {lev.x, lev.y, lev.z, lev.w};
    for (int q = 0; q < 4; q++) {
        R.le1[4 * (size_t)i + q] = g[q];
        lb.ev[4 * (size_t)i + q] = lev_a[q] + g[4 + q];
        lb.ea[4 * (size_t)i + q] = hdt * lev_a[q];
    }
    const double lv_a[3] = {lv.x, lv.y, lv.z}, lacc_a[3] = {lacc.x, lacc.y, lacc.z};
    for (int q = 0; q < 3; q++) {
        lb.vel[3 * (size_t)i + q] = lv_a[q];
        lb.acc[3 * (size_t)i + q] = hdt * lv_a[q];
        R.bar[10 * (size_t)i + q] = inv_m * (hdt * lv_a[q] + lacc_a[q]);
        R.bar[10 * (size_t)i + 3 + q] = g[8 + q];
        R.bar[10 * (size_t)i + 6 + q] = g[11 + q];
        R.lom[3 * (size_t)i + q] = 0.0;
        R.lpos1[3 * (size_t)i + q] = 0.0;
    }
    R.bar[10 * (size_t)i + 9] = g[14];
}

// One sphere-sphere contact differentiated with respect to x_i (0-2), x_j (3-5), v_i (6-8),
// v_j (9-11), w_i (12-14), w_j (15-17), the history entry read (18-19), theta (20-22); the
// previous pair's tangential force enters as a constant.
constexpr int kPairN = 23;
// One wall contact: x_i (0-2), v_i (3-5), w_i (6-8), history (9-10), theta (11-13).
constexpr int kWallN = 14;

__device__ inline void contacts_vjp(const RevArgs& R, uint32_t i, double (&gth)[3])
{
    const CalArgs& A = R.A;
    const StepRec& r = R.r;
    const Theta<double> thd = {A.mu, A.rest, A.s};
    const V4<double> ip = ld4d(r.pos1 + 4 * (size_t)i);
    const V3<double> xi = {ip.x, ip.y, ip.z};
    const double ir = ip.w, im = A.mass[i];
    const V3<double> vi = ld3d(r.vel + 3 * (size_t)i);
    const V3<double> wi = ang_vel(ld4d(r.ep1 + 4 * (size_t)i), ld4d(r.ev + 4 * (size_t)i));
    const double* bar = R.bar + 10 * (size_t)i;
    const V3<double> bF = {bar[0], bar[1], bar[2]}, bM = {bar[3], bar[4], bar[5]}, bT = {bar[6], bar[7], bar[8]};
    const double bR = bar[9];
    const uint32_t nt = r.tape_n[i];
    const size_t sb = (size_t)i * A.cap;
    V3<double> Ft = {0.0, 0.0, 0.0};   // the previous pair's tangential force (values)
    V3<double> gx = {0.0, 0.0, 0.0}, gv = {0.0, 0.0, 0.0}, gw = {0.0, 0.0, 0.0};
    uint32_t written_pp = 0u, written_pl = 0u;   // slots this step wrote (bit masks)
    uint32_t m = 0;
    for (; m < nt; m++) {
        const uint32_t code = r.tape_code[(size_t)i * kTape + m];
        if (code & kPlaneTag) break;
        const uint16_t sl = r.tape_slot[(size_t)i * kTape + m];
        const int slot_in = (int)(sl & 0xffu) - 1, slot_out = (int)(sl >> 8) - 1;
        double sd0 = 0.0, sd1 = 0.0;
        if (slot_in >= 0) {
            sd0 = r.pp_tsd[2 * (sb + slot_in)];
            sd1 = r.pp_tsd[2 * (sb + slot_in) + 1];
        }
        const uint32_t k = code;
        const V4<double> jp = ld4d(r.pos1 + 4 * (size_t)k);
        const V3<double> xj = {jp.x, jp.y, jp.z};
        const V3<double> vj = ld3d(r.vel + 3 * (size_t)k);
        const V3<double> wj = ang_vel(ld4d(r.ep1 + 4 * (size_t)k), ld4d(r.ev + 4 * (size_t)k));
        // cotangents of this pair's new history entry
        double bs0 = 0.0, bs1 = 0.0;
        if (slot_out >= 0) {
            bs0 = R.la.pp_tsd[2 * (sb + slot_out)];
            bs1 = R.la.pp_tsd[2 * (sb + slot_out) + 1];
            written_pp |= 1u << slot_out;
        }
        {
            typedef Dual<kPairN> D;
            const Theta<D> th = {D::input(A.mu, 20), D::input(A.rest, 21), D::input(A.s, 22)};
            const PairK<D> c = pair_k<D>(A.law.model, A.mat_pp, th, ir, jp.w, im, A.mass[k]);
            D s0 = D::input(sd0, 18), s1 = D::input(sd1, 19);
            V3<D> FtD = const3<kPairN>(Ft);
            Sums<D> acc = zero_sums<D>();
            pair_side<D>(A.law, c, A.mat_pp.rfric, seed3<kPairN>(xi, 0), ir, seed3<kPairN>(xj, 3), jp.w,
                         seed3<kPairN>(vi, 6), seed3<kPairN>(vj, 9), seed3<kPairN>(wi, 12), seed3<kPairN>(wj, 15), s0,
                         s1, FtD, acc);
            double g[kPairN];
            for (int q = 0; q < kPairN; q++) g[q] = 0.0;
            vjp_add3(g, acc.F, bF);
            vjp_add3(g, acc.M, bM);
            vjp_add3(g, acc.T, bT);
            vjp_add(g, acc.R, bR);
            vjp_add(g, s0, bs0);
            vjp_add(g, s1, bs1);
            gx = gx + V3<double>{g[0], g[1], g[2]};
            atomic_add3(R.lpos1 + 3 * (size_t)k, g[3], g[4], g[5]);
            gv = gv + V3<double>{g[6], g[7], g[8]};
            atomic_add3(R.lb.vel + 3 * (size_t)k, g[9], g[10], g[11]);
            gw = gw + V3<double>{g[12], g[13], g[14]};
            atomic_add3(R.lom + 3 * (size_t)k, g[15], g[16], g[17]);
            if (slot_in >= 0) {
                R.lb.pp_tsd[2 * (sb + slot_in)] += g[18];
                R.lb.pp_tsd[2 * (sb + slot_in) + 1] += g[19];
            }
            gth[0] += g[20];
            gth[1] += g[21];
            gth[2] += g[22];
        }
        // the value pass, for the next pair's stale tangential force
        {
            const PairK<double> c = pair_k<double>(A.law.model, A.mat_pp, thd, ir, jp.w, im, A.mass[k]);
            Sums<double> acc = zero_sums<double>();
            double s0 = sd0, s1 = sd1;
            pair_side<double>(A.law, c, A.mat_pp.rfric, xi, ir, xj, jp.w, vi, vj, wi, wj, s0, s1, Ft, acc);
        }
    }
    const size_t lb = (size_t)i * kPlSlots;
    for (uint32_t w = 0; w < A.nplane && m < nt; w++) {
        if (r.tape_code[(size_t)i * kTape + m] != (kPlaneTag | w)) continue;
        const uint16_t sl = r.tape_slot[(size_t)i * kTape + m];
        const int slot_in = (int)(sl & 0xffu) - 1, slot_out = (int)(sl >> 8) - 1;
        m++;
        double sd0 = 0.0, sd1 = 0.0;
        if (slot_in >= 0) {
            sd0 = r.pl_tsd[2 * (lb + slot_in)];
            sd1 = r.pl_tsd[2 * (lb + slot_in) + 1];
        }
        double bs0 = 0.0, bs1 = 0.0;
        if (slot_out >= 0) {
            bs0 = R.la.pl_tsd[2 * (lb + slot_out)];
            bs1 = R.la.pl_tsd[2 * (lb + slot_out) + 1];
            written_pl |= 1u << slot_out;
        }
        typedef Dual<kWallN> D;
        const Theta<D> th = {D::input(A.mu, 11), D::input(A.rest, 12), D::input(A.s, 13)};
        const PairK<D> c = pair_k<D>(A.law.model, A.mat_pl[w], th, ir, 0.0, im, 0.0);
        D s0 = D::input(sd0, 9), s1 = D::input(sd1, 10);
        Sums<D> acc = zero_sums<D>();
        plane_side<D>(A.law, c, A.mat_pl[w].rfric, A.planes[w], A.plane_dbi + 15 * (size_t)w, seed3<kWallN>(xi, 0), ir,
                      seed3<kWallN>(vi, 3), seed3<kWallN>(wi, 6), s0, s1, acc);
        double g[kWallN];
        for (int q = 0; q < kWallN; q++) g[q] = 0.0;
        vjp_add3(g, acc.F, bF);
        vjp_add3(g, acc.M, bM);
        vjp_add3(g, acc.T, bT);
        vjp_add(g, acc.R, bR);
        vjp_add(g, s0, bs0);
        vjp_add(g, s1, bs1);
        gx = gx + V3<double>{g[0], g[1], g[2]};
        gv = gv + V3<double>{g[3], g[4], g[5]};
        gw = gw + V3<double>{g[6], g[7], g[8]};
        if (slot_in >= 0) {
            R.lb.pl_tsd[2 * (lb + slot_in)] += g[9];
            R.lb.pl_tsd[2 * (lb + slot_in) + 1] += g[10];
        }
        gth[0] += g[11];
        gth[1] += g[12];
        gth[2] += g[13];
    }
    atomic_add3(R.lpos1 + 3 * (size_t)i, gx.x, gx.y, gx.z);
    atomic_add3(R.lb.vel + 3 * (size_t)i, gv.x, gv.y, gv.z);
    atomic_add3(R.lom + 3 * (size_t)i, gw.x, gw.y, gw.z);
    // slots this step did not write keep their values: their cotangents pass through
    for (uint32_t q = 0; q < A.cap; q++)
        if (!(written_pp & (1u << q))) {
            R.lb.pp_tsd[2 * (sb + q)] += R.la.pp_tsd[2 * (sb + q)];
            R.lb.pp_tsd[2 * (sb + q) + 1] += R.la.pp_tsd[2 * (sb + q) + 1];
        }
    for (int q = 0; q < kPlSlots; q++)
        if (!(written_pl & (1u << q))) {
            R.lb.pl_tsd[2 * (lb + q)] += R.la.pl_tsd[2 * (lb + q)];
            R.lb.pl_tsd[2 * (lb + q) + 1] += R.la.pl_tsd[2 * (lb + q) + 1];
        }
}

__global__ __launch_bounds__(64) void k_cal_rev_contacts(RevArgs R)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double gth[3] = {0.0, 0.0, 0.0};
    if (i < R.A.n) contacts_vjp(R, i, gth);
    // theta: a wave sum (every lane takes part), then one atomic per wave and component
    for (int q = 0; q < 3; q++) {
        double v = gth[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63u) == 0u && v != 0.0) atomicAdd(R.gtheta + q, v);
    }
}

// The angular velocity w(e', e_dot) and the position update of sphere i: the cotangents of
// the state before the step.
constexpr int kHeadN = 12;   // e (0-3), e_dot (4-7), e_ddot (8-11)

__global__ __launch_bounds__(64) void k_cal_rev_head(RevArgs R)
{
    const CalArgs& A = R.A;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const StepRec& r = R.r;
    const Lam &la = R.la, &lb = R.lb;
    // w = 2 G(e') e_dot: its cotangent adds to e' and e_dot
    double le1[4], lev[4];
    for (int q = 0; q < 4; q++) {
        le1[q] = R.le1[4 * (size_t)i + q] + la.ep[4 * (size_t)i + q];
        lev[q] = lb.ev[4 * (size_t)i + q];
    }
    {
        typedef Dual<8> D;
        const V3<D> w = ang_vel(seed4<8>(ld4d(r.ep1 + 4 * (size_t)i), 0), seed4<8>(ld4d(r.ev + 4 * (size_t)i), 4));
        double g[8];
        for (int q = 0; q < 8; q++) g[q] = 0.0;
        vjp_add3(g, w, ld3d(R.lom + 3 * (size_t)i));
        for (int q = 0; q < 4; q++) {
            le1[q] += g[q];
            lev[q] += g[4 + q];
        }
    }
    // position: p' = p + dt v + h a
    const V3<double> lp = ld3d(la.pos + 3 * (size_t)i) + ld3d(R.lpos1 + 3 * (size_t)i);
    for (int q = 0; q < 3; q++) {
        const double l = q == 0 ? lp.x : (q == 1 ? lp.y : lp.z);
        lb.pos[3 * (size_t)i + q] = l;
        lb.vel[3 * (size_t)i + q] += A.dt * l;
        lb.acc[3 * (size_t)i + q] += A.h * l;
    }
    // Euler parameters: e' = normalise(e + dt e_dot + h e_ddot)
    typedef Dual<kHeadN> D;
    V3<D> pd = {D(0.0), D(0.0), D(0.0)};
    V4<D> e = seed4<kHeadN>(ld4d(r.ep0 + 4 * (size_t)i), 0);
    position_update<D>(A.dt, A.h, pd, e, V3<D>{D(0.0), D(0.0), D(0.0)}, seed4<kHeadN>(ld4d(r.ev + 4 * (size_t)i), 4),
                       V3<D>{D(0.0), D(0.0), D(0.0)}, seed4<kHeadN>(ld4d(r.ea + 4 * (size_t)i), 8));
    double g[kHeadN];
    for (int q = 0; q < kHeadN; q++) g[q] = 0.0;
    vjp_add(g, e.x, le1[0]);
    vjp_add(g, e.y, le1[1]);
    vjp_add(g, e.z, le1[2]);
    vjp_add(g, e.w, le1[3]);
    for (int q = 0; q < 4; q++) {
        lb.ep[4 * (size_t)i + q] = g[q];
        lb.ev[4 * (size_t)i + q] = lev[q] + g[4 + q];
        lb.ea[4 * (size_t)i + q] += g[8 + q];
    }
}

__global__ void k_fill_zero(double* p, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = 0.0;
}

}  // namespace

// ============================================================================ host side
struct xdem_calib {
    xdem_calib_config cfg;
    hipStream_t stream = nullptr;
    uint32_t n = 0, cap = 12, nplane = 0, ncell = 0, every = 100, max_steps = 0, nseg = 0;
    double dt = 0.0, h = 0.0, g[3] = {0, 0, 0}, wo[3] = {0, 0, 0}, cs = 0.0;
    uint32_t gsize[3] = {128, 128, 128};
    Law law;
    Material mat_pp;
    std::vector<double> h_pos, h_vel, h_acc;   // initial state
    std::vector<void*> allocs;
    // device
    double *mass = nullptr, *inertia = nullptr;
    Material* d_mat_pl = nullptr;
    xdem_plane_info* d_planes = nullptr;
    double* d_plane_dbi = nullptr;
    StateView st;                               // current state (vel/ev: buffer 0 or 1)
    double *vel_buf[2] = {nullptr, nullptr}, *ev_buf[2] = {nullptr, nullptr};
    int sv = 0;
    uint32_t *hash = nullptr, *index = nullptr, *sorted = nullptr, *cstart = nullptr, *cend = nullptr;
    uint32_t* err = nullptr;
    // checkpoints: nseg states (doubles and slot arrays, packed)
    char* ckpt = nullptr;
    size_t state_bytes = 0;
    // segment records: every step's StepRec arrays and tape, packed per step
    char* recs = nullptr;
    size_t rec_bytes = 0;
    // reverse sweep
    double* lam[2] = {nullptr, nullptr};        // two Lam sets, packed
    size_t lam_doubles = 0;
    double *le1 = nullptr, *lom = nullptr, *lpos1 = nullptr, *bar = nullptr, *gtheta = nullptr;
    double theta[3] = {0, 0, 0};
    unsigned first_step = 0, last_steps = 0;   // the steps of the last forward run at theta
    bool have_forward = false;

    template <class P> hipError_t alloc(P** p, size_t bytes)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes ? bytes : 16);
        if (e != hipSuccess) return e;
        allocs.push_back(q);
        *p = static_cast<P*>(q);
        return hipSuccess;
    }

    // packed layout of a state: pos 4n, vel 3n, acc 3n, ep 4n, ev 4n, ea 4n, tmax 3n, rres n,
    // pp_tsd 2 cap n, pl_tsd 6n (doubles), then pp_count n, pp_id cap n, pl_count n, pl_id 3n
    size_t state_doubles() const { return (size_t)n * (4 + 3 + 3 + 4 + 4 + 4 + 3 + 1 + 2 * cap + 2 * kPlSlots); }
    size_t state_u32() const { return (size_t)n * (1 + cap + 1 + kPlSlots); }

    CalArgs args() const
    {
        CalArgs A;
        std::memset(&A, 0, sizeof(A));
        A.n = n; A.cap = cap; A.nplane = nplane; A.ncell = ncell;
        A.dt = dt; A.h = h; A.gx = g[0]; A.gy = g[1]; A.gz = g[2];
        for (int k = 0; k < 3; k++) { A.wo[k] = wo[k]; A.gsize[k] = gsize[k]; }
        A.cs = cs;
        A.law = law;
        A.flat_inertia = cfg.quirks.flat_inertia;
        A.mat_pp = mat_pp;
        A.mat_pl = d_mat_pl;
        A.planes = d_planes;
        A.plane_dbi = d_plane_dbi;
        A.mu = theta[0]; A.rest = theta[1]; A.s = theta[2];
        A.mass = mass;
        A.inertia = inertia;
        A.st = st;
        A.st.vel = vel_buf[sv];
        A.st.ev = ev_buf[sv];
        A.vel_out = vel_buf[sv ^ 1];
        A.ev_out = ev_buf[sv ^ 1];
        A.hash = hash; A.index = index; A.sorted = sorted; A.cstart = cstart; A.cend = cend;
        A.err = err;
        return A;
    }

    // copy the current state to/from a packed buffer
    hipError_t pack(char* dst) const
    {
        double* d = reinterpret_cast<double*>(dst);
        uint32_t* u = reinterpret_cast<uint32_t*>(d + state_doubles());
        const size_t N = n;
        const double* src[10] = {st.pos, vel_buf[sv], st.acc, st.ep, ev_buf[sv], st.ea, st.tmax, st.rres, st.pp_tsd, st.pl_tsd};
        const size_t len[10] = {4 * N, 3 * N, 3 * N, 4 * N, 4 * N, 4 * N, 3 * N, N, 2 * cap * N, 2 * kPlSlots * N};
        size_t o = 0;
        hipError_t e = hipSuccess;
        for (int k = 0; k < 10 && e == hipSuccess; k++) {
            e = hipMemcpyAsync(d + o, src[k], sizeof(double) * len[k], hipMemcpyDeviceToDevice, stream);
            o += len[k];
        }
        const uint32_t* us[4] = {st.pp_count, st.pp_id, st.pl_count, st.pl_id};
        const size_t ul[4] = {N, cap * N, N, kPlSlots * N};
        o = 0;
        for (int k = 0; k < 4 && e == hipSuccess; k++) {
            e = hipMemcpyAsync(u + o, us[k], sizeof(uint32_t) * ul[k], hipMemcpyDeviceToDevice, stream);
            o += ul[k];
        }
        return e;
    }
    hipError_t unpack(const char* srcb)
    {
        const double* d = reinterpret_cast<const double*>(srcb);
        const uint32_t* u = reinterpret_cast<const uint32_t*>(d + state_doubles());
        sv = 0;
        const size_t N = n;
        double* dst[10] = {st.pos, vel_buf[0], st.acc, st.ep, ev_buf[0], st.ea, st.tmax, st.rres, st.pp_tsd, st.pl_tsd};
        const size_t len[10] = {4 * N, 3 * N, 3 * N, 4 * N, 4 * N, 4 * N, 3 * N, N, 2 * cap * N, 2 * kPlSlots * N};
        size_t o = 0;
        hipError_t e = hipSuccess;
        for (int k = 0; k < 10 && e == hipSuccess; k++) {
            e = hipMemcpyAsync(dst[k], d + o, sizeof(double) * len[k], hipMemcpyDeviceToDevice, stream);
            o += len[k];
        }
        uint32_t* ud[4] = {st.pp_count, st.pp_id, st.pl_count, st.pl_id};
        const size_t ul[4] = {N, cap * N, N, kPlSlots * N};
        o = 0;
        for (int k = 0; k < 4 && e == hipSuccess; k++) {
            e = hipMemcpyAsync(ud[k], u + o, sizeof(uint32_t) * ul[k], hipMemcpyDeviceToDevice, stream);
            o += ul[k];
        }
        return e;
    }

    // record layout of one step: pos1 4n, ep0 4n, ep1 4n, vel 3n, ev 4n, acc 3n, ea 4n, tmax 3n,
    // rres n, pp_tsd 2 cap n, pl_tsd 6n (doubles); pp_count n, pp_id cap n, pl_count n, pl_id 3n,
    // tape_code kTape n (u32); tape_slot kTape n (u16); tape_n n (u8)
    size_t rec_doubles() const { return (size_t)n * (4 + 4 + 4 + 3 + 4 + 3 + 4 + 3 + 1 + 2 * cap + 2 * kPlSlots); }
    size_t rec_size() const
    {
        size_t b = sizeof(double) * rec_doubles() + sizeof(uint32_t) * ((size_t)n * (1 + cap + 1 + kPlSlots + kTape));
        b += sizeof(uint16_t) * (size_t)n * kTape + (size_t)n;
        return (b + 255) & ~(size_t)255;
    }
    StepRec rec_view(size_t k) const
    {
        char* base = recs + k * rec_bytes;
        double* d = reinterpret_cast<double*>(base);
        const size_t N = n;
        StepRec r;
        size_t o = 0;
        r.pos1 = d + o; o += 4 * N;
        r.ep0 = d + o; o += 4 * N;
        r.ep1 = d + o; o += 4 * N;
        r.vel = d + o; o += 3 * N;
        r.ev = d + o; o += 4 * N;
        r.acc = d + o; o += 3 * N;
        r.ea = d + o; o += 4 * N;
        r.tmax = d + o; o += 3 * N;
        r.rres = d + o; o += N;
        r.pp_tsd = d + o; o += 2 * cap * N;
        r.pl_tsd = d + o; o += 2 * kPlSlots * N;
        uint32_t* u = reinterpret_cast<uint32_t*>(d + o);
        o = 0;
        r.pp_count = u + o; o += N;
        r.pp_id = u + o; o += cap * N;
        r.pl_count = u + o; o += N;
        r.pl_id = u + o; o += kPlSlots * N;
        r.tape_code = u + o; o += kTape * N;
        uint16_t* s16 = reinterpret_cast<uint16_t*>(u + o);
        r.tape_slot = s16;
        r.tape_n = reinterpret_cast<const uint8_t*>(s16 + kTape * N);
        return r;
    }

    Lam lam_view(int k) const
    {
        double* p = lam[k];
        const size_t N = n;
        Lam L;
        L.pos = p; p += 3 * N;
        L.vel = p; p += 3 * N;
        L.acc = p; p += 3 * N;
        L.ep = p; p += 4 * N;
        L.ev = p; p += 4 * N;
        L.ea = p; p += 4 * N;
        L.pp_tsd = p; p += 2 * cap * N;
        L.pl_tsd = p; p += 2 * kPlSlots * N;
        return L;
    }

    // one forward step; rec (optional) receives the step's record
    hipError_t step(const StepRec* rec, size_t rec_index)
    {
        hipError_t e;
        const size_t N = n;
        const dim3 grid((n + kThreads - 1) / kThreads);
        char* base = rec ? recs + rec_index * rec_bytes : nullptr;
        if (rec) {
            // the state before the step: ep (before the update), vel, ev, acc, ea, tmax, rres, slots
            double* d = reinterpret_cast<double*>(base);
            const size_t o_ep0 = 4 * N, o_vel = 12 * N, o_ev = 15 * N, o_acc = 19 * N, o_ea = 22 * N, o_tmax = 26 * N,
                         o_rres = 29 * N, o_pp = 30 * N, o_pl = 30 * N + 2 * cap * N;
            if ((e = hipMemcpyAsync(d + o_ep0, st.ep, sizeof(double) * 4 * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(d + o_vel, vel_buf[sv], sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(d + o_ev, ev_buf[sv], sizeof(double) * 4 * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(d + o_acc, st.acc, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(d + o_ea, st.ea, sizeof(double) * 4 * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(d + o_tmax, st.tmax, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(d + o_rres, st.rres, sizeof(double) * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(d + o_pp, st.pp_tsd, sizeof(double) * 2 * cap * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(d + o_pl, st.pl_tsd, sizeof(double) * 2 * kPlSlots * N, hipMemcpyDeviceToDevice, stream))) return e;
            uint32_t* u = reinterpret_cast<uint32_t*>(d + rec_doubles());
            if ((e = hipMemcpyAsync(u, st.pp_count, sizeof(uint32_t) * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(u + N, st.pp_id, sizeof(uint32_t) * cap * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(u + N + cap * N, st.pl_count, sizeof(uint32_t) * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(u + 2 * N + cap * N, st.pl_id, sizeof(uint32_t) * kPlSlots * N, hipMemcpyDeviceToDevice, stream))) return e;
        }
        CalArgs A = args();
        hipLaunchKernelGGL(k_cal_position, grid, dim3(kThreads), 0, stream, A);
        if ((e = hipGetLastError())) return e;
        if (rec) {
            double* d = reinterpret_cast<double*>(base);
            if ((e = hipMemcpyAsync(d, st.pos, sizeof(double) * 4 * N, hipMemcpyDeviceToDevice, stream))) return e;
            if ((e = hipMemcpyAsync(d + 8 * N, st.ep, sizeof(double) * 4 * N, hipMemcpyDeviceToDevice, stream))) return e;
        }
        // the launcher layer's sort runs on the null stream
        if ((e = hipStreamSynchronize(stream))) return e;
        const int rc = xdem_cu_reorderDataAndFindCellStart(hash, index, cstart, cend, sorted, n, ncell);
        if (rc != XDEM_OK) return hipErrorUnknown;
        if ((e = hipDeviceSynchronize())) return e;
        if (rec) {
            const StepRec r = rec_view(rec_index);
            A.tape_code = const_cast<uint32_t*>(r.tape_code);
            A.tape_slot = const_cast<uint16_t*>(r.tape_slot);
            A.tape_n = const_cast<uint8_t*>(r.tape_n);
        }
        hipLaunchKernelGGL(k_cal_contacts, grid, dim3(kThreads), 0, stream, A);
        if ((e = hipGetLastError())) return e;
        sv ^= 1;
        return hipSuccess;
    }

    hipError_t reset()
    {
        hipError_t e;
        const size_t N = n;
        sv = 0;
        if ((e = hipMemcpyAsync(st.pos, h_pos.data(), sizeof(double) * 4 * N, hipMemcpyHostToDevice, stream))) return e;
        if ((e = hipMemcpyAsync(vel_buf[0], h_vel.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice, stream))) return e;
        if ((e = hipMemcpyAsync(st.acc, h_acc.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice, stream))) return e;
        std::vector<double> ep(4 * N, 0.0);
        for (size_t i = 0; i < N; i++) ep[4 * i] = 1.0;
        if ((e = hipMemcpyAsync(st.ep, ep.data(), sizeof(double) * 4 * N, hipMemcpyHostToDevice, stream))) return e;
        if ((e = hipStreamSynchronize(stream))) return e;   // ep is a local
        if ((e = hipMemsetAsync(ev_buf[0], 0, sizeof(double) * 4 * N, stream))) return e;
        if ((e = hipMemsetAsync(st.ea, 0, sizeof(double) * 4 * N, stream))) return e;
        if ((e = hipMemsetAsync(st.tmax, 0, sizeof(double) * 3 * N, stream))) return e;
        if ((e = hipMemsetAsync(st.rres, 0, sizeof(double) * N, stream))) return e;
        if ((e = hipMemsetAsync(st.pp_count, 0, sizeof(uint32_t) * N, stream))) return e;
        if ((e = hipMemsetAsync(st.pp_id, 0, sizeof(uint32_t) * cap * N, stream))) return e;
        if ((e = hipMemsetAsync(st.pp_tsd, 0, sizeof(double) * 2 * cap * N, stream))) return e;
        if ((e = hipMemsetAsync(st.pl_count, 0, sizeof(uint32_t) * N, stream))) return e;
        if ((e = hipMemsetAsync(st.pl_id, 0, sizeof(uint32_t) * kPlSlots * N, stream))) return e;
        if ((e = hipMemsetAsync(st.pl_tsd, 0, sizeof(double) * 2 * kPlSlots * N, stream))) return e;
        return hipSuccess;
    }
};

#define CAL_CHECK(x)                                 \
    do {                                             \
        hipError_t e_ = (x);                         \
        if (e_ != hipSuccess) return XDEM_E_HIP;     \
    } while (0)

extern "C" {

int xdem_calib_create(const xdem_scene* s, const xdem_calib_config* cfg_in, xdem_calib** out)
{
    if (!s || !out || s->np == 0 || !s->pos || !s->mass || !s->inertia) return XDEM_E_INVALID;
    if (s->nplane && (!s->planes || !s->plane_dbi || !s->plane_cp)) return XDEM_E_INVALID;
    if (!(s->dt > 0.0) || !(s->cell_size > 0.0)) return XDEM_E_INVALID;
    if (s->contact_model != 0 && s->contact_model != 1) return XDEM_E_INVALID;
    xdem_calib_config cfg = {};
    if (cfg_in) cfg = *cfg_in;
    if (cfg.quirks.velocity_not_updated) return XDEM_E_UNSUPPORTED;
    if (s->cp_pp.coh != 0.0) return XDEM_E_UNSUPPORTED;
    for (unsigned k = 0; k < s->nplane; k++)
        if (s->plane_cp[k].coh != 0.0 || s->planes[k].id != k) return XDEM_E_UNSUPPORTED;
    const unsigned cap = cfg.capacity ? cfg.capacity : 12u;
    if (cap > (unsigned)kCapMax || s->nplane > 8u) return XDEM_E_INVALID;
    xdem_calib* c = new (std::nothrow) xdem_calib();
    if (!c) return XDEM_E_INVALID;
    c->cfg = cfg;
    c->n = s->np;
    c->cap = cap;
    c->nplane = s->nplane;
    c->every = cfg.checkpoint_every ? cfg.checkpoint_every : 100u;
    c->max_steps = cfg.max_steps ? cfg.max_steps : c->every;
    c->nseg = (c->max_steps + c->every - 1) / c->every;
    c->dt = s->dt;
    c->h = 0.5 * s->dt * s->dt;
    for (int k = 0; k < 3; k++) { c->g[k] = s->gravity[k]; c->wo[k] = s->world_origin[k]; }
    c->cs = s->cell_size;
    c->ncell = c->gsize[0] * c->gsize[1] * c->gsize[2];
    c->law.model = s->contact_model;
    c->law.gpu_law = cfg.quirks.gpu_force_law != 0;
    c->law.stale_ft = cfg.quirks.stale_ft != 0;
    c->law.dt = s->dt;
    const xdem_contact_property& p = s->cp_pp;
    c->mat_pp = {p.Ei, p.Ej, p.pri, p.prj, p.rfric, p.sratio};
    const size_t N = s->np;
    c->h_pos.assign(s->pos, s->pos + 4 * N);
    c->h_vel.assign(3 * N, 0.0);
    if (s->vel) c->h_vel.assign(s->vel, s->vel + 3 * N);
    c->h_acc.assign(3 * N, 0.0);
    for (size_t i = 0; i < N; i++) c->h_acc[3 * i + 1] = s->gravity[1];   // acc = (0, g.y, 0) at the start
    auto fail = [&](int rc) {
        xdem_calib_destroy(c);
        return rc;
    };
    if (hipSetDevice(cfg.device) != hipSuccess) return fail(XDEM_E_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(XDEM_E_HIP);
    if (xdem_reserve_workspace(s->np) != XDEM_OK) return fail(XDEM_E_HIP);
#define ALLOC(p, bytes) if (c->alloc(&(p), (bytes)) != hipSuccess) return fail(XDEM_E_HIP)
    ALLOC(c->mass, sizeof(double) * N);
    ALLOC(c->inertia, sizeof(double) * 3 * N);
    ALLOC(c->d_mat_pl, sizeof(Material) * std::max(1u, s->nplane));
    ALLOC(c->d_planes, sizeof(xdem_plane_info) * std::max(1u, s->nplane));
    ALLOC(c->d_plane_dbi, sizeof(double) * 15 * std::max(1u, s->nplane));
    ALLOC(c->st.pos, sizeof(double) * 4 * N);
    ALLOC(c->vel_buf[0], sizeof(double) * 3 * N);
    ALLOC(c->vel_buf[1], sizeof(double) * 3 * N);
    ALLOC(c->st.acc, sizeof(double) * 3 * N);
    ALLOC(c->st.ep, sizeof(double) * 4 * N);
    ALLOC(c->ev_buf[0], sizeof(double) * 4 * N);
    ALLOC(c->ev_buf[1], sizeof(double) * 4 * N);
    ALLOC(c->st.ea, sizeof(double) * 4 * N);
    ALLOC(c->st.tmax, sizeof(double) * 3 * N);
    ALLOC(c->st.rres, sizeof(double) * N);
    ALLOC(c->st.pp_count, sizeof(uint32_t) * N);
    ALLOC(c->st.pp_id, sizeof(uint32_t) * cap * N);
    ALLOC(c->st.pp_tsd, sizeof(double) * 2 * cap * N);
    ALLOC(c->st.pl_count, sizeof(uint32_t) * N);
    ALLOC(c->st.pl_id, sizeof(uint32_t) * kPlSlots * N);
    ALLOC(c->st.pl_tsd, sizeof(double) * 2 * kPlSlots * N);
    ALLOC(c->hash, sizeof(uint32_t) * N);
    ALLOC(c->index, sizeof(uint32_t) * N);
    ALLOC(c->sorted, sizeof(uint32_t) * N);
    ALLOC(c->cstart, sizeof(uint32_t) * c->ncell);
    ALLOC(c->cend, sizeof(uint32_t) * c->ncell);
    ALLOC(c->err, sizeof(uint32_t));
    c->state_bytes = (sizeof(double) * c->state_doubles() + sizeof(uint32_t) * c->state_u32() + 255) & ~(size_t)255;
    ALLOC(c->ckpt, c->state_bytes * c->nseg);
    c->rec_bytes = c->rec_size();
    ALLOC(c->recs, c->rec_bytes * c->every);
    c->lam_doubles = N * (3 + 3 + 3 + 4 + 4 + 4 + 2 * (size_t)cap + 2 * kPlSlots);
    ALLOC(c->lam[0], sizeof(double) * c->lam_doubles);
    ALLOC(c->lam[1], sizeof(double) * c->lam_doubles);
    ALLOC(c->le1, sizeof(double) * 4 * N);
    ALLOC(c->lom, sizeof(double) * 3 * N);
    ALLOC(c->lpos1, sizeof(double) * 3 * N);
    ALLOC(c->bar, sizeof(double) * 10 * N);
    ALLOC(c->gtheta, sizeof(double) * 3);
#undef ALLOC
    std::vector<Material> mpl(std::max(1u, s->nplane));
    for (unsigned k = 0; k < s->nplane; k++) {
        const xdem_contact_property& q = s->plane_cp[k];
        mpl[k] = {q.Ei, q.Ej, q.pri, q.prj, q.rfric, q.sratio};
    }
    hipStream_t st = c->stream;
    if (hipMemcpyAsync(c->mass, s->mass, sizeof(double) * N, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(c->inertia, s->inertia, sizeof(double) * 3 * N, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(c->d_mat_pl, mpl.data(), sizeof(Material) * mpl.size(), hipMemcpyHostToDevice, st) ||
        (s->nplane && hipMemcpyAsync(c->d_planes, s->planes, sizeof(xdem_plane_info) * s->nplane, hipMemcpyHostToDevice, st)) ||
        (s->nplane && hipMemcpyAsync(c->d_plane_dbi, s->plane_dbi, sizeof(double) * 15 * s->nplane, hipMemcpyHostToDevice, st)) ||
        hipMemsetAsync(c->err, 0, sizeof(uint32_t), st) || hipStreamSynchronize(st))
        return fail(XDEM_E_HIP);
    *out = c;
    return XDEM_OK;
}

int xdem_calib_destroy(xdem_calib* c)
{
    if (!c) return XDEM_E_INVALID;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return XDEM_OK;
}

int xdem_calib_forward(xdem_calib* c, const double theta[3], unsigned int nsteps, double* pos_out)
{
    if (!c || !theta) return XDEM_E_INVALID;
    if (nsteps > c->max_steps) return XDEM_E_INVALID;
    if (!(theta[0] >= 0.0) || !(theta[1] > 0.0) || !std::isfinite(theta[2])) return XDEM_E_INVALID;
    CAL_CHECK(hipSetDevice(c->cfg.device));
    for (int k = 0; k < 3; k++) c->theta[k] = theta[k];
    CAL_CHECK(c->reset());
    for (unsigned k = 0; k < nsteps; k++) {
        if (k % c->every == 0) CAL_CHECK(c->pack(c->ckpt + (k / c->every) * c->state_bytes));
        CAL_CHECK(c->step(nullptr, 0));
    }
    CAL_CHECK(hipStreamSynchronize(c->stream));
    c->first_step = 0;
    c->last_steps = nsteps;
    c->have_forward = true;
    if (pos_out)
        CAL_CHECK(hipMemcpy(pos_out, c->st.pos, sizeof(double) * 4 * (size_t)c->n, hipMemcpyDeviceToHost));
    return XDEM_OK;
}

int xdem_calib_forward_from(xdem_calib* c, const double theta[3], unsigned int start, unsigned int nsteps,
                            double* pos_out)
{
    if (!c || !theta) return XDEM_E_INVALID;
    if (!c->have_forward) return XDEM_E_NOT_READY;
    // the run restarts from a stored state of the last run: at a checkpoint it passed
    if (start % c->every != 0 || start > c->last_steps || start >= nsteps || nsteps > c->max_steps) return XDEM_E_INVALID;
    if (!(theta[0] >= 0.0) || !(theta[1] > 0.0) || !std::isfinite(theta[2])) return XDEM_E_INVALID;
    CAL_CHECK(hipSetDevice(c->cfg.device));
    if (start == c->last_steps) {
        // the end state of the last run is not a stored one unless a checkpoint was written there
        CAL_CHECK(c->pack(c->ckpt + (start / c->every) * c->state_bytes));
    }
    CAL_CHECK(c->unpack(c->ckpt + (start / c->every) * c->state_bytes));
    for (int k = 0; k < 3; k++) c->theta[k] = theta[k];
    for (unsigned k = start; k < nsteps; k++) {
        if (k % c->every == 0 && k != start) CAL_CHECK(c->pack(c->ckpt + (k / c->every) * c->state_bytes));
        CAL_CHECK(c->step(nullptr, 0));
    }
    CAL_CHECK(hipStreamSynchronize(c->stream));
    c->first_step = start;
    c->last_steps = nsteps;
    if (pos_out)
        CAL_CHECK(hipMemcpy(pos_out, c->st.pos, sizeof(double) * 4 * (size_t)c->n, hipMemcpyDeviceToHost));
    return XDEM_OK;
}

int xdem_calib_state(xdem_calib* c, double* pos, double* vel, double* ep, double* ev)
{
    if (!c) return XDEM_E_INVALID;
    CAL_CHECK(hipSetDevice(c->cfg.device));
    CAL_CHECK(hipStreamSynchronize(c->stream));
    const size_t N = c->n;
    if (pos) CAL_CHECK(hipMemcpy(pos, c->st.pos, sizeof(double) * 4 * N, hipMemcpyDeviceToHost));
    if (vel) CAL_CHECK(hipMemcpy(vel, c->vel_buf[c->sv], sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
    if (ep) CAL_CHECK(hipMemcpy(ep, c->st.ep, sizeof(double) * 4 * N, hipMemcpyDeviceToHost));
    if (ev) CAL_CHECK(hipMemcpy(ev, c->ev_buf[c->sv], sizeof(double) * 4 * N, hipMemcpyDeviceToHost));
    return XDEM_OK;
}

int xdem_calib_backward(xdem_calib* c, const double* dL_dpos, double grad[3])
{
    if (!c || !dL_dpos || !grad) return XDEM_E_INVALID;
    if (!c->have_forward) return XDEM_E_NOT_READY;
    CAL_CHECK(hipSetDevice(c->cfg.device));
    hipStream_t st = c->stream;
    const size_t N = c->n;
    const dim3 zg(256), zb(256);
    hipLaunchKernelGGL(k_fill_zero, zg, zb, 0, st, c->lam[0], c->lam_doubles);
    hipLaunchKernelGGL(k_fill_zero, zg, zb, 0, st, c->gtheta, (size_t)3);
    CAL_CHECK(hipGetLastError());
    int la = 0;
    CAL_CHECK(hipMemcpyAsync(c->lam_view(la).pos, dL_dpos, sizeof(double) * 3 * N, hipMemcpyHostToDevice, st));
    const unsigned nsteps = c->last_steps;
    const dim3 grid((c->n + 63) / 64), block(64);
    const int seg0 = (int)(c->first_step / c->every);
    for (int seg = (int)((nsteps + c->every - 1) / c->every) - 1; seg >= seg0; seg--) {
        const unsigned k0 = (unsigned)seg * c->every, k1 = std::min(nsteps, k0 + c->every);
        // recompute the segment from its checkpoint, recording every step
        CAL_CHECK(c->unpack(c->ckpt + (size_t)seg * c->state_bytes));
        for (unsigned k = k0; k < k1; k++) {
            const StepRec r = c->rec_view(k - k0);
            CAL_CHECK(c->step(&r, k - k0));
        }
        // and sweep it backwards
        for (int k = (int)k1 - 1; k >= (int)k0; k--) {
            RevArgs R;
            std::memset(&R, 0, sizeof(R));
            R.A = c->args();
            R.r = c->rec_view((size_t)(k - (int)k0));
            R.la = c->lam_view(la);
            R.lb = c->lam_view(la ^ 1);
            R.le1 = c->le1; R.lom = c->lom; R.lpos1 = c->lpos1; R.bar = c->bar; R.gtheta = c->gtheta;
            hipLaunchKernelGGL(k_fill_zero, zg, zb, 0, st, R.lb.pp_tsd, (size_t)(2 * c->cap + 2 * kPlSlots) * N);
            hipLaunchKernelGGL(k_cal_rev_tail, grid, block, 0, st, R);
            hipLaunchKernelGGL(k_cal_rev_contacts, grid, block, 0, st, R);
            hipLaunchKernelGGL(k_cal_rev_head, grid, block, 0, st, R);
            CAL_CHECK(hipGetLastError());
            la ^= 1;
        }
    }
    CAL_CHECK(hipMemcpyAsync(grad, c->gtheta, sizeof(double) * 3, hipMemcpyDeviceToHost, st));
    CAL_CHECK(hipStreamSynchronize(st));
    return XDEM_OK;
}

int xdem_calib_error_word(xdem_calib* c, unsigned int* out, int clear)
{
    if (!c || !out) return XDEM_E_INVALID;
    CAL_CHECK(hipSetDevice(c->cfg.device));
    CAL_CHECK(hipStreamSynchronize(c->stream));
    CAL_CHECK(hipMemcpy(out, c->err, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (clear) CAL_CHECK(hipMemset(c->err, 0, sizeof(uint32_t)));
    return XDEM_OK;
}

int xdem_calib_pair_eval(const xdem_contact_property* cp, int contact_model, const xdem_quirks* q, double dt,
                         double ir, double jr, double im, double jm, const double in[23], double out[12], double* jac)
{
    if (!cp || !q || !in || !out) return XDEM_E_INVALID;
    Law L;
    L.model = contact_model;
    L.gpu_law = q->gpu_force_law != 0;
    L.stale_ft = q->stale_ft != 0;
    L.dt = dt;
    const Material m = {cp->Ei, cp->Ej, cp->pri, cp->prj, cp->rfric, cp->sratio};
    for (int k = 0; k < 12; k++) out[k] = 0.0;
    if (jac)
        for (int k = 0; k < 12 * kPairN; k++) jac[k] = 0.0;
    typedef Dual<kPairN> D;
    const Theta<D> th = {D::input(in[20], 20), D::input(in[21], 21), D::input(in[22], 22)};
    const PairK<D> c = pair_k<D>(contact_model, m, th, ir, jr, im, jm);
    auto v3 = [&](int k) { return V3<double>{in[k], in[k + 1], in[k + 2]}; };
    D s0 = D::input(in[18], 18), s1 = D::input(in[19], 19);
    V3<D> Ft = {D(0.0), D(0.0), D(0.0)};
    Sums<D> acc = zero_sums<D>();
    const bool hit = pair_side<D>(L, c, cp->rfric, seed3<kPairN>(v3(0), 0), ir, seed3<kPairN>(v3(3), 3), jr,
                                  seed3<kPairN>(v3(6), 6), seed3<kPairN>(v3(9), 9), seed3<kPairN>(v3(12), 12),
                                  seed3<kPairN>(v3(15), 15), s0, s1, Ft, acc);
    if (!hit) return 0;
    const D* o[12] = {&acc.F.x, &acc.F.y, &acc.F.z, &acc.M.x, &acc.M.y, &acc.M.z, &acc.T.x, &acc.T.y, &acc.T.z, &acc.R, &s0, &s1};
    for (int r = 0; r < 12; r++) {
        out[r] = o[r]->v;
        if (jac)
            for (int k = 0; k < kPairN; k++) jac[r * kPairN + k] = o[r]->d[k];
    }
    return 1;
}

}  // extern "C"

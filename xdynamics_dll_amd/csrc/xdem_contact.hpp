// xdem_contact.hpp -- one-contact evaluations of the DEM step, shared by the launcher
// kernels (reference array layouts) and the fused step kernels.
//
// FP64 instantiations keep the reference's operation order (impl:613-733 for spheres,
// impl:950-1080 for walls, impl:1748-1835 for triangles). FP32 instantiations (the mixed
// path) evaluate the tangential spring-damper magnitude and the moment arithmetic in
// float; the pair test, positions, relative velocity, normal force, tangential history
// and all sums stay in double (see force_law).
#pragma once

#include <type_traits>

#include "xdem_dev.hpp"

namespace xdem {

// Material constants of one contact object, computed once (host or device).
__host__ __device__ inline ContactConst make_contact_const(const xdem_contact_property& p, double reach_radius)
{
    const double pi = 3.14159265358979323846;
    ContactConst c;
    const double iE = p.Ei, jE = p.Ej, ip = p.pri, jp = p.prj;
    const double Eeq = (iE * jE) / (iE * (1 - jp * jp) + jE * (1 - ip * ip));
    double Seq = (2.0 * (2.0 - ip) * (1.0 + ip) / iE) + (2.0 * (2.0 - jp) * (1.0 + jp) / jE);
    Seq = 1.0 / Seq;
    const double lne = log(p.rest);
    const double beta = -lne * sqrt(1.0 / (lne * lne + pi * pi));
    c.k43E = (4.0 / 3.0) * Eeq;
    c.beta = beta;
    c.ks_hm = 8.0 * Seq;
    c.vs_hm = 2.0 * sqrt(5.0 / 6.0) * beta;
    c.vs_dhs = beta * p.sratio;
    c.sratio = p.sratio;
    c.fric = p.fric;
    c.rfric = p.rfric;
    c.coh = p.coh;
    const double Eeq_c = 1.0 / (((1.0 - p.pri * p.pri) / p.Ei) + ((1.0 - p.prj * p.prj) / p.Ej));
    c.c3cp = 3.0 * p.coh * pi;
    c.E4 = 4.0 * Eeq_c;
    c.c8 = 8.0 * pi * p.coh * Eeq_c;
    c.pcc = pi * pi * p.coh * p.coh;
    c.ee = Eeq_c * Eeq_c;
    c.jk = -(3.0 / 2.0) * pi * p.coh;
    c.coh_reach = 0.0;
    if (p.coh && reach_radius > 0.0) {
        // |depth| = 0.75 (pcc R*/ee)^(1/3) grows with R*; bound it at the largest R* of
        // the contact object and widen by a relative margin for the rounding of pow.
        c.coh_reach = 0.75 * cbrt((c.pcc * reach_radius) / c.ee) * (1.0 + 1e-6);
    }
    return c;
}

// Accumulators of one sphere over all its contacts (always FP64).
struct Accum {
    V3<double> F, M, T;  // force, moment, rolling torque sum (tmax)
    double R;            // rolling resistance magnitude (rres)
};

// -------------------------------------------------------------------- sphere-sphere
// Relative geometry of a candidate pair: rp = pj - pi, dist = |rp|, cdist = ri + rj - dist.
// The predicate cdist > 0 is evaluated in FP64 without contraction (bit-exact pair sets).
struct PairGeom {
    V3<double> rp;
    double dist, cdist;
};

__device__ __forceinline__ PairGeom pair_geom(V4<double> pi_, V4<double> pj)
{
#pragma clang fp contract(off)
    PairGeom g;
    g.rp = mk3(pj.x - pi_.x, pj.y - pi_.y, pj.z - pi_.z);
    g.dist = sqrt(g.rp.x * g.rp.x + g.rp.y * g.rp.y + g.rp.z * g.rp.z);
    g.cdist = (pi_.w + pj.w) - g.dist;
    return g;
}

// Contact of sphere i with sphere j (cdist > 0), impl:679-706. Updates the i-side sums
// and the pair's history (sd0, sd1). Ft carries over between calls (quirk Q4) unless
// the caller resets it.
template <class T, int MODEL, bool GPU_LAW>
__device__ __forceinline__ void sphere_contact(const ContactConst& cc, double dt, V4<double> ipos, V3<double> ivel,
                                               V3<double> iom, double im, V4<double> jpos, V3<double> jvel,
                                               V3<double> jom, double jm, const PairGeom& g, double& sd0, double& sd1,
                                               V3<T>& Ft, Accum& acc)
{
    const double ir = ipos.w, jr = jpos.w;
    const PairConst<T> c = pair_const<T, MODEL>(cc, ir, jr, im, jm);
    const V3<double> unit = g.rp / g.dist;
    const V3<double> ip3 = mk3(ipos.x, ipos.y, ipos.z);
    const V3<double> cpt = ip3 + ir * unit;
    const V3<double> dcpr = cpt - ip3;
    const V3<double> dcpr_j = cpt - mk3(jpos.x, jpos.y, jpos.z);
    const V3<double> rv = (jvel + cross(jom, dcpr_j)) - (ivel + cross(iom, dcpr));
    V3<T> Fn;
    force_law<T, MODEL, GPU_LAW>(cc, c, g.cdist, dt, sd0, sd1, rv, unit, Ft, Fn);
    const V3<T> fnt = Fn + Ft;
    const V3<T> dc = cvt3<T>(dcpr);
    acc.T = acc.T + cvt3<double>(cross(dc, fnt));
    acc.R += (double)(c.eq_r * T(cc.rfric) * length(Fn));
    acc.F = acc.F + cvt3<double>(fnt);
    acc.M = acc.M + cvt3<double>(cross(dc, fnt));
}

// One contact of sphere i with sphere j evaluated on its own (the pair pass): the parts of
// sphere_contact that do not depend on the earlier pairs of the sphere. out = Fn (3), Ft (3),
// the lever arm dc (3), in T and stored as double, and the rolling-resistance term; returns
// whether Ft was set (|e| != 0). sphere_pair_sum then adds the pairs in visiting order with
// the same operations as sphere_contact, carrying Ft over where it was not set (quirk Q4).
template <class T, int MODEL, bool GPU_LAW>
__device__ __forceinline__ bool sphere_pair_eval(const ContactConst& cc, double dt, V4<double> ipos, V3<double> ivel,
                                                 V3<double> iom, double im, V4<double> jpos, V3<double> jvel,
                                                 V3<double> jom, double jm, const PairGeom& g, double& sd0,
                                                 double& sd1, double out[10])
{
    const double ir = ipos.w, jr = jpos.w;
    const PairConst<T> c = pair_const<T, MODEL>(cc, ir, jr, im, jm);
    const V3<double> unit = g.rp / g.dist;
    const V3<double> ip3 = mk3(ipos.x, ipos.y, ipos.z);
    const V3<double> cpt = ip3 + ir * unit;
    const V3<double> dcpr = cpt - ip3;
    const V3<double> dcpr_j = cpt - mk3(jpos.x, jpos.y, jpos.z);
    const V3<double> rv = (jvel + cross(jom, dcpr_j)) - (ivel + cross(iom, dcpr));
    V3<T> Fn, Ft = mk3(T(0), T(0), T(0));
    const bool set = force_law<T, MODEL, GPU_LAW>(cc, c, g.cdist, dt, sd0, sd1, rv, unit, Ft, Fn);
    const V3<T> dc = cvt3<T>(dcpr);
    out[0] = (double)Fn.x; out[1] = (double)Fn.y; out[2] = (double)Fn.z;
    out[3] = (double)Ft.x; out[4] = (double)Ft.y; out[5] = (double)Ft.z;
    out[6] = (double)dc.x; out[7] = (double)dc.y; out[8] = (double)dc.z;
    out[9] = (double)(c.eq_r * T(cc.rfric) * length(Fn));
    return set;
}

// Adds one evaluated contact to the sphere's sums: sphere_contact's accumulation, in its order.
template <class T>
__device__ __forceinline__ void sphere_pair_sum(const double v[10], bool set, bool stale_ft, V3<T>& Ft, Accum& acc)
{
    if (!stale_ft) Ft = mk3(T(0), T(0), T(0));
    if (set) Ft = mk3(T(v[3]), T(v[4]), T(v[5]));
    const V3<T> Fn = mk3(T(v[0]), T(v[1]), T(v[2]));
    const V3<T> dc = mk3(T(v[6]), T(v[7]), T(v[8]));
    const V3<T> fnt = Fn + Ft;
    acc.T = acc.T + cvt3<double>(cross(dc, fnt));
    acc.R += v[9];
    acc.F = acc.F + cvt3<double>(fnt);
    acc.M = acc.M + cvt3<double>(cross(dc, fnt));
}

// Cohesion-only attraction of a separated pair (impl:708-718), FP64 for every precision.
__device__ __forceinline__ void sphere_cohesion(const ContactConst& cc, double ir, double jr, const PairGeom& g, Accum& acc)
{
    const double Req = jr ? (ir * jr) / (ir + jr) : ir;
    const double f = cc.jk * Req;
    const double cf = cohesion_force<double>(cc, Req, f);
    const V3<double> unit = g.rp / g.dist;
    acc.F = acc.F - cf * unit;
}

// -------------------------------------------------------------------------- walls
// One sphere against one wall (impl:995-1066). `slots` is the sphere's wall history
// (ids and (ds, s_dot)) in registers; the wall's id selects reset-or-append (quirk Q8).
struct WallSlots {
    uint32_t id[XDEM_MAX_P2PL_COUNT];
    double s0[XDEM_MAX_P2PL_COUNT], s1[XDEM_MAX_P2PL_COUNT];
    uint32_t count;
    uint32_t dirty;              // a slot was written (the caller writes the slots back only then)
    uint32_t loaded;             // slots read from memory (only done once a wall needs them)
    const uint32_t* src_id;      // the sphere's slots in memory
    const double* src_tsd;       // (ds, s_dot_prev) pairs

    __device__ __forceinline__ void init(uint32_t cnt, const uint32_t* ids, const double* tsd)
    {
        count = cnt; dirty = 0; loaded = 0; src_id = ids; src_tsd = tsd;
    }
    __device__ __forceinline__ void load()
    {
        if (loaded) return;
#pragma unroll
        for (int s = 0; s < XDEM_MAX_P2PL_COUNT; s++) {
            id[s] = src_id[s];
            s0[s] = src_tsd[2 * s];
            s1[s] = src_tsd[2 * s + 1];
        }
        loaded = 1;
    }
};

// True when the sphere cannot touch the wall, for a wall without cohesion, even grown by
// `skin` (a sphere that has moved less than the skin stays out of reach). Every contact
// region of plane_detection needs (dp . uw)^2 < r^2 max(1, |uw|^2): corners and edges test
// (dp . uw)^2 plus non-negative terms against r^2, the face needs |dp . uw| / |uw| < r.
// Beyond that, with a 1e-6 margin for rounding, the outcome is the no-contact one.
__device__ __forceinline__ bool wall_out_of_reach_raw(double coh, V3<double> xw, V3<double> uw, double uw2,
                                                      V4<double> ipos, double skin)
{
    if (coh != 0.0) return false;
    const V3<double> dp = mk3(ipos.x, ipos.y, ipos.z) - xw;
    const double wz = dot(dp, uw);
    const double r = ipos.w + skin;
    return wz * wz >= (r * r) * fmax(1.0, uw2) * (1.0 + 1e-6);
}

__device__ __forceinline__ bool wall_out_of_reach(const ContactConst& cc, const xdem_plane_info& pl, V4<double> ipos,
                                                  double skin = 0.0)
{
    const double uw2 = pl.uw.x * pl.uw.x + pl.uw.y * pl.uw.y + pl.uw.z * pl.uw.z;
    return wall_out_of_reach_raw(cc.coh, mk3(pl.xw.x, pl.xw.y, pl.xw.z), mk3(pl.uw.x, pl.uw.y, pl.uw.z), uw2, ipos,
                                 skin);
}

// The state update of wall_contact for a sphere out of the wall's reach: no force, the
// slot count reset by the first wall (id 0), as impl:989 does.
__device__ __forceinline__ void wall_no_contact(const xdem_plane_info& pl, WallSlots& ws, Accum& acc)
{
    uint32_t n_new = pl.id ? ws.count : 0u;
    if (n_new > XDEM_MAX_P2PL_COUNT) n_new = XDEM_MAX_P2PL_COUNT;
    ws.count = n_new;
    acc.F = acc.F + mk3(0.0, 0.0, 0.0);
    acc.M = acc.M + mk3(0.0, 0.0, 0.0);
    acc.T = acc.T + mk3(0.0, 0.0, 0.0);
    acc.R += 0.0;
}

template <class T, int MODEL, bool GPU_LAW>
__device__ __forceinline__ bool wall_contact(const ContactConst& cc, const xdem_plane_info& pl, const double* dbi, double dt,
                                             V4<double> ipos, V3<double> ivel, V3<double> iom, double m,
                                             WallSlots& ws, Accum& acc, double bf[6])
{
    const V3<double> ip3 = mk3(ipos.x, ipos.y, ipos.z);
    const double r = ipos.w;
    const uint32_t k = pl.id;
    uint32_t n_new = k ? ws.count : 0u;
    if (n_new > XDEM_MAX_P2PL_COUNT) n_new = XDEM_MAX_P2PL_COUNT;
    const V3<double> dp = ip3 - mk3(pl.xw.x, pl.xw.y, pl.xw.z);
    V3<double> unit = mk3(0.0, 0.0, 0.0);
    const V3<double> wp = mk3(dot(dp, mk3(pl.u1.x, pl.u1.y, pl.u1.z)), dot(dp, mk3(pl.u2.x, pl.u2.y, pl.u2.z)),
                              dot(dp, mk3(pl.uw.x, pl.uw.y, pl.uw.z)));
    if (wall_out_of_reach(cc, pl, ipos)) {
        wall_no_contact(pl, ws, acc);
        return false;
    }
    const double cdist = plane_detection(pl, ip3, wp, unit, r);
    double coh_s = 0;
    if (cc.coh) coh_s = cohesion_limit_depth(cc, r);
    double sd0 = 0.0, sd1 = 0.0;
    const V3<double> cpt = ip3 + r * unit;
    const V3<double> bpos = mk3(dbi[XDEM_PX], dbi[XDEM_PY], dbi[XDEM_PZ]);
    const V3<double> dcpr_j = cpt - bpos;
    V3<double> sumF = mk3(0.0, 0.0, 0.0), sumM = mk3(0.0, 0.0, 0.0), tma = mk3(0.0, 0.0, 0.0);
    double res = 0.0;
    bool stored = false, overflow = false;
    if (cdist > 0) {
        ws.load();
        // the first slot with this wall's id (constant indices: the slots stay in registers)
        bool found = false;
#pragma unroll
        for (int s = 0; s < XDEM_MAX_P2PL_COUNT; s++)
            if (!found && ws.id[s] == k) { sd0 = ws.s0[s]; sd1 = ws.s1[s]; found = true; }
        const V3<double> dcpr = cpt - ip3;
        V4<double> be = {dbi[XDEM_E0], dbi[XDEM_E1], dbi[XDEM_E2], dbi[XDEM_E3]};
        V4<double> bd = {dbi[XDEM_ED0], dbi[XDEM_ED1], dbi[XDEM_ED2], dbi[XDEM_ED3]};
        const V3<double> oj = omega_of(be, bd);
        const V3<double> bvel = mk3(dbi[XDEM_VX], dbi[XDEM_VY], dbi[XDEM_VZ]);
        const PairConst<T> c = pair_const<T, MODEL>(cc, r, 0.0, m, 0.0);
        V3<T> Fn, Ft = mk3(T(0), T(0), T(0));
        const V3<double> dv = (bvel + cross(oj, dcpr_j)) - (ivel + cross(iom, dcpr));
        force_law<T, MODEL, GPU_LAW>(cc, c, cdist, dt, sd0, sd1, dv, unit, Ft, Fn);
        const V3<double> mf = cvt3<double>(Fn + Ft);
        tma = tma + cross(dcpr, mf);
        res += (double)(c.eq_r * T(cc.rfric) * length(Fn));
        sumF = sumF + mf;
        sumM = sumM + cross(dcpr, mf);
        bf[0] += -mf.x; bf[1] += -mf.y; bf[2] += -mf.z;
        const V3<double> bm = -cross(dcpr_j, mf);
        bf[3] += bm.x; bf[4] += bm.y; bf[5] += bm.z;
        stored = true;
    } else if (cdist <= 0 && fabs(cdist) < fabs(coh_s)) {
        const double f = cc.jk * r;
        const double cf = cohesion_force<double>(cc, r, f);
        const V3<double> mf = (-cf) * unit;
        sumF = sumF + mf;
        bf[0] += -mf.x; bf[1] += -mf.y; bf[2] += -mf.z;
        const V3<double> bm = -cross(dcpr_j, mf);
        bf[3] += bm.x; bf[4] += bm.y; bf[5] += bm.z;
        stored = true;
    }
    if (stored) {
        if (n_new < XDEM_MAX_P2PL_COUNT) {
            ws.load();
#pragma unroll
            for (int s = 0; s < XDEM_MAX_P2PL_COUNT; s++)
                if ((uint32_t)s == n_new) { ws.id[s] = k; ws.s0[s] = sd0; ws.s1[s] = sd1; }
            ws.dirty = 1;
            n_new++;
        } else {
            overflow = true;
        }
    }
    ws.count = n_new;
    acc.F = acc.F + sumF;
    acc.M = acc.M + sumM;
    acc.T = acc.T + tma;
    acc.R += res;
    return overflow;
}

// ---------------------------------------------------------------------- triangles
struct TriHit {
    uint32_t id;
    V3<double> cpt;
};

// One sphere against one triangle contact point (impl:1748-1835). Returns true when a
// history entry (id, sd0, sd1) is to be stored.
template <class T, int MODEL, bool GPU_LAW>
__device__ __forceinline__ bool triangle_contact(const ContactConst& cc, const double* dbi, double dt, TriHit h,
                                                 V3<double> ipos, double r, double m, V3<double> ivel, V3<double> iom,
                                                 const uint32_t* old_id, const double* old_s0, const double* old_s1,
                                                 double& sd0, double& sd1, Accum& acc, double bf[6])
{
    const V3<double> rp = h.cpt - ipos;
    sd0 = 0.0; sd1 = 0.0;
    const V3<double> po2cp = h.cpt - mk3(dbi[XDEM_PX], dbi[XDEM_PY], dbi[XDEM_PZ]);
    const double dist = length(rp);
    const V3<double> unit = rp / dist;
    const double cdist = r - dist;
    double coh_s = 0;
    if (cc.coh) coh_s = cohesion_limit_depth(cc, r);
    if (cdist <= 0 && fabs(cdist) < fabs(coh_s)) {
        const double f = cc.jk * r;
        const double cf = cohesion_force<double>(cc, r, f);
        acc.F = acc.F - cf * unit;
        return true;
    }
    V4<double> be = {dbi[XDEM_E0], dbi[XDEM_E1], dbi[XDEM_E2], dbi[XDEM_E3]};
    V4<double> bd = {dbi[XDEM_ED0], dbi[XDEM_ED1], dbi[XDEM_ED2], dbi[XDEM_ED3]};
    const V3<double> jom = omega_of(be, bd);
    for (uint32_t s = 0; s < XDEM_MAX_P2MS_COUNT; s++)
        if (old_id[s] == h.id) { sd0 = old_s0[s]; sd1 = old_s1[s]; break; }
    const V3<double> rc = h.cpt - ipos;
    const V3<double> bvel = mk3(dbi[XDEM_VX], dbi[XDEM_VY], dbi[XDEM_VZ]);
    const V3<double> dv = (bvel + cross(jom, po2cp)) - (ivel + cross(iom, rc));
    const PairConst<T> c = pair_const<T, MODEL>(cc, r, 0.0, m, dbi[XDEM_MM]);
    V3<T> Fn, Ft = mk3(T(0), T(0), T(0));
    force_law<T, MODEL, GPU_LAW>(cc, c, cdist, dt, sd0, sd1, dv, unit, Ft, Fn);
    const V3<double> mf = cvt3<double>(Fn + Ft);
    acc.T = acc.T + cross(rc, mf);
    acc.R += (double)(c.eq_r * T(cc.rfric) * length(Fn));
    acc.F = acc.F + mf;
    acc.M = acc.M + cross(rc, mf);
    bf[0] += -mf.x; bf[1] += -mf.y; bf[2] += -mf.z;
    const V3<double> mm = -cross(po2cp, mf);
    bf[3] += mm.x; bf[4] += mm.y; bf[5] += mm.z;
    return true;
}

// ---------------------------------------------------------------------- cylinders
// Ends of the cylinder's axis at the body pose: pos + A(e) pbase and pos + A(e) ptop.
struct CylAxis {
    V3<double> pos, base, top, ab;
    V4<double> e;
};

__device__ __forceinline__ CylAxis cylinder_axis(const xdem_cylinder_info& cy, const double* bi)
{
    CylAxis a;
    a.pos = mk3(bi[XDEM_PX], bi[XDEM_PY], bi[XDEM_PZ]);
    a.e = {bi[XDEM_E0], bi[XDEM_E1], bi[XDEM_E2], bi[XDEM_E3]};
    a.base = a.pos + rot_global(mk3(cy.pbase.x, cy.pbase.y, cy.pbase.z), a.e);
    a.top = a.pos + rot_global(mk3(cy.ptop.x, cy.ptop.y, cy.ptop.z), a.e);
    a.ab = a.top - a.base;
    return a;
}

// Overlap of a sphere (centre p, radius r) with the cylinder (impl:1189-1288); u is the unit
// vector along which the force acts, inner is set for a sphere inside the tube (or on its
// axis). The regions, in the reference's order:
//  - beside the tube (axial parameter t in [0, 1]): the side wall at len_rr.y, from outside
//    (dist > len_rr.z) or from inside (u reversed); a sphere on the axis returns 0;
//  - beyond an end, within len_rr.z (+ thickness; > len_rr.z when there is a thickness) of
//    the axis: the lid plane, at half the length from the body position, u from the axis
//    point towards the body position;
//  - otherwise the rim circle of radius len_rr.z at the nearer end (the top unless the local
//    y offset from it exceeds the length), its point in the local xz plane found by the
//    atan / sin / cos construction of the kernel text: overlap r - dist when positive, -1
//    otherwise (r - dist in any case with rim_gap, for the reach test's gap).
__device__ __forceinline__ double cylinder_detection(const xdem_cylinder_info& cy, const CylAxis& a, V3<double> p,
                                                     double r, V3<double>& u, bool& inner, bool rim_gap = false)
{
    const double t = dot(p - a.base, a.ab) / dot(a.ab, a.ab);
    const V3<double> cp = a.base + t * a.ab;
    const double dist = length(p - cp);
    if (t >= 0 && t <= 1) {
        if (dist == 0) {
            inner = true;
            return 0;
        }
        u = (cp - p) / dist;
        if (dist < cy.len_rr.z) {
            inner = true;
            u = -u;
            return dist + r - cy.len_rr.y;
        }
        return cy.len_rr.y + r - dist;
    }
    const double thick = cy.thickness;
    const double one = thick ? 1.0 : 0.0;
    if (dist < cy.len_rr.z + thick && dist > one * cy.len_rr.z) {
        const V3<double> oc = a.pos - cp;
        const double ocl = length(oc);
        u = oc / ocl;
        return cy.len_rr.x * 0.5 + r - ocl;
    }
    V3<double> at = rot_local(p - a.top, a.e);
    V3<double> rim = a.top;
    if (fabs(at.y) > cy.len_rr.x) {
        at = rot_local(p - a.base, a.e);
        rim = a.base;
    }
    const double rr = cy.len_rr.z;
    const double phi = atan(at.x / at.z);
    double lx;
    if (phi < 0 && at.z < 0) lx = rr * sin(-phi);
    else if (phi > 0 && at.x < 0 && at.z < 0) lx = rr * sin(-phi);
    else lx = rr * sin(phi);
    double lz = rr * cos(phi);
    if (at.z < 0 && lz > 0) lz = -lz;
    else if (at.z > 0 && lz < 0) lz = -lz;
    const V3<double> c = rim + rot_global(mk3(lx, 0.0, lz), a.e);
    const V3<double> dv = c - p;
    const double d = length(dv);
    u = dv / d;
    if (d < r || rim_gap) return r - d;
    return -1.0;
}

// The closed lids seen from inside the tube (impl:1290-1325): an open lid (empty_part 1: top
// for t > 0.6, 2: bottom for t < 0.4) gives 0 with u and cpt untouched; within len_rr.z of
// the axis, u points from the body position along the axis towards the sphere, the contact
// point is p + r u and the overlap r + |axial offset| - length / 2; -1 otherwise. With
// open_lids false both lids count as closed (the reach test: a sphere leaving through an open
// end meets the lid region beyond it, whose overlap continues this one's).
__device__ __forceinline__ double cylinder_inner_lid(const xdem_cylinder_info& cy, const CylAxis& a, V3<double> p,
                                                     double r, V3<double>& u, V3<double>& cpt, bool open_lids = true)
{
    const double t = dot(p - a.base, a.ab) / dot(a.ab, a.ab);
    const V3<double> cp = a.base + t * a.ab;
    const double dist = length(p - cp);
    if (open_lids && cy.empty_part == 1 && t > 0.6) return 0.0;
    if (open_lids && cy.empty_part == 2 && t < 0.4) return 0.0;
    double gab = -1.0;
    if (dist < cy.len_rr.z) {
        const V3<double> oc = a.pos - cp;
        const double ocl = length(oc);
        u = (-oc) / ocl;
        cpt = p + r * u;
        gab = r + ocl - cy.len_rr.x * 0.5;
    }
    return gab;
}

// A sphere's cylinder history slots: the previous step's entries (looked up over all slots,
// as the kernel does) and the entries of this step appended after `base`.
struct CylSlots {
    uint32_t old_id[XDEM_MAX_P2CY_COUNT];
    double old_s0[XDEM_MAX_P2CY_COUNT], old_s1[XDEM_MAX_P2CY_COUNT];
    uint32_t id[XDEM_MAX_P2CY_COUNT];
    double s0[XDEM_MAX_P2CY_COUNT], s1[XDEM_MAX_P2CY_COUNT];
    uint32_t count;      // entries valid after the call
    bool overflow;
};

// One contact of the sphere with the cylinder body, or its cohesive attraction: force,
// moment, rolling sums, body wrench and a history entry with id pid (impl:1111-1187 and the
// cohesion branches of impl:1391-1460).
template <class T, int MODEL, bool GPU_LAW>
__device__ __forceinline__ void cylinder_pair(const ContactConst& cc, const double* bi, double dt, V3<double> ip3,
                                              double r, double m, V3<double> ivel, V3<double> iom, V3<double> cpt,
                                              V3<double> unit, double cdist, double coh_s, uint32_t pid,
                                              CylSlots& cs, Accum& acc, double bf[6])
{
    const V3<double> bpos = mk3(bi[XDEM_PX], bi[XDEM_PY], bi[XDEM_PZ]);
    double sd0 = 0.0, sd1 = 0.0;
    bool stored = false;
    if (cdist > 0) {
        const V3<double> po2cp = cpt - bpos;
        V4<double> be = {bi[XDEM_E0], bi[XDEM_E1], bi[XDEM_E2], bi[XDEM_E3]};
        V4<double> bd = {bi[XDEM_ED0], bi[XDEM_ED1], bi[XDEM_ED2], bi[XDEM_ED3]};
        const V3<double> om = omega_of(be, bd);
        for (uint32_t s = 0; s < XDEM_MAX_P2CY_COUNT; s++)
            if (cs.old_id[s] == pid) { sd0 = cs.old_s0[s]; sd1 = cs.old_s1[s]; break; }
        const V3<double> rc = cpt - ip3;
        const V3<double> dv = mk3(bi[XDEM_VX], bi[XDEM_VY], bi[XDEM_VZ]) + cross(om, po2cp) - (ivel + cross(iom, rc));
        const PairConst<T> c = pair_const<T, MODEL>(cc, r, 0.0, m, 0.0);
        V3<T> Fn, Ft = mk3(T(0), T(0), T(0));
        force_law<T, MODEL, GPU_LAW>(cc, c, cdist, dt, sd0, sd1, dv, unit, Ft, Fn);
        const V3<double> mf = cvt3<double>(Fn + Ft);
        acc.T = acc.T + cross(rc, mf);
        acc.R += (double)(c.eq_r * T(cc.rfric) * length(Fn));
        acc.F = acc.F + mf;
        acc.M = acc.M + cross(rc, mf);
        bf[0] += -mf.x; bf[1] += -mf.y; bf[2] += -mf.z;
        const V3<double> bm = -cross(po2cp, mf);
        bf[3] += bm.x; bf[4] += bm.y; bf[5] += bm.z;
        stored = true;
    } else if (cdist <= 0 && fabs(cdist) < fabs(coh_s)) {
        const double f = cc.jk * r;
        const double cf = cohesion_force<double>(cc, r, f);
        const V3<double> mf = (-cf) * unit;
        acc.F = acc.F + mf;
        bf[0] += -mf.x; bf[1] += -mf.y; bf[2] += -mf.z;
        const V3<double> bm = -cross(cpt - bpos, mf);
        bf[3] += bm.x; bf[4] += bm.y; bf[5] += bm.z;
        stored = true;
    }
    if (!stored) return;
    if (cs.count < XDEM_MAX_P2CY_COUNT) {
        for (uint32_t s = 0; s < XDEM_MAX_P2CY_COUNT; s++)
            if (s == cs.count) { cs.id[s] = pid; cs.s0[s] = sd0; cs.s1[s] = sd1; }
        cs.count++;
    } else {
        cs.overflow = true;
    }
}

// The sphere against one cylinder body (cylinder_contact_force_kernel, impl:1327-1460): the
// first cylinder (id 0) starts the slots afresh, a later one appends after the entries
// already there. cs.id/s0/s1 must hold the previous entries (copied from cs.old_*).
template <class T, int MODEL, bool GPU_LAW>
__device__ __forceinline__ void cylinder_contact(const ContactConst& cc, const xdem_cylinder_info& cy, const double* bi,
                                                 double dt, V4<double> ipos, V3<double> ivel, V3<double> iom, double m,
                                                 CylSlots& cs, Accum& acc, double bf[6])
{
    const V3<double> ip3 = mk3(ipos.x, ipos.y, ipos.z);
    const double r = ipos.w;
    const uint32_t k = cy.id;
    if (!k) cs.count = 0;
    else if (cs.count > XDEM_MAX_P2CY_COUNT) cs.count = XDEM_MAX_P2CY_COUNT;
    const CylAxis a = cylinder_axis(cy, bi);
    bool inner = false;
    V3<double> unit = mk3(0.0, 0.0, 0.0);
    double cdist = cylinder_detection(cy, a, ip3, r, unit, inner);
    double coh_s = 0;
    if (cc.coh) coh_s = cohesion_limit_depth(cc, r);
    V3<double> cpt = ip3 + r * unit;
    Accum part = {mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), mk3(0.0, 0.0, 0.0), 0.0};
    cylinder_pair<T, MODEL, GPU_LAW>(cc, bi, dt, ip3, r, m, ivel, iom, cpt, unit, cdist, coh_s, k, cs, part, bf);
    if (inner) {
        cdist = cylinder_inner_lid(cy, a, ip3, r, unit, cpt);
        cylinder_pair<T, MODEL, GPU_LAW>(cc, bi, dt, ip3, r, m, ivel, iom, cpt, unit, cdist, coh_s, k + 1000u, cs,
                                         part, bf);
    }
    acc.F = acc.F + part.F;
    acc.M = acc.M + part.M;
    acc.T = acc.T + part.T;
    acc.R += part.R;
}

// An upper bound of the sphere's cylinder overlap at radius r, negative by the distance it
// lacks: the larger of the two detections, the lids taken as closed, the rim's overlap signed. In each region the overlap is r plus a function of the
// centre that changes by at most the distance the centre moves; where regions meet, the
// functions differ by at most the cylinder's margin (host-computed from its geometry). So a
// sphere whose bound with its radius grown by skin + margin is not positive cannot touch
// before it or the body has moved skin / 2.
__device__ __forceinline__ double cylinder_reach(const xdem_cylinder_info& cy, const CylAxis& a, V3<double> p, double r)
{
    bool inner = false;
    V3<double> u = mk3(0.0, 0.0, 0.0), cpt = p;
    double g = cylinder_detection(cy, a, p, r, u, inner, true);
    if (inner) g = fmax(g, cylinder_inner_lid(cy, a, p, r, u, cpt, false));
    return g;
}

// True when the sphere cannot touch the cylinder even grown by `skin` (any cohesive cylinder:
// never, as a cohesive pair attracts within its separation depth and the reference stores a
// slot even at zero overlap).
__device__ __forceinline__ bool cylinder_out_of_reach(const ContactConst& cc, const xdem_cylinder_info& cy,
                                                      const double* dbi, double margin, V4<double> p, double skin)
{
    if (cc.coh != 0.0) return false;
    const CylAxis a = cylinder_axis(cy, dbi);
    const double g = cylinder_reach(cy, a, mk3(p.x, p.y, p.z), (p.w + skin + margin) * (1.0 + 1e-9) + 1e-12);
    return !(g > 0.0);
}

// ---------------------------------------------------------------- rolling resistance
// impl:2420-2454 on the accumulated sums; returns true when tmax/rres are consumed
// (reset), false when rres == 0 and they are left as they are.
__device__ __forceinline__ bool rolling_moment(double dt, double J, V3<double> omega, V3<double> Tm, double Tr, V3<double>& M)
{
    if (!Tr) return false;
    const double Jdt = J * dt;
    const V3<double> T = mk3(Jdt * omega.x - Tm.x, Jdt * omega.y - Tm.y, Jdt * omega.z - Tm.z);
    const double lT = length(T);
    if (lT && Tr) {
        V3<double> Trv = Tr * (T / lT);
        if (length(Trv) >= lT) Trv = T;
        M = M + Trv;
    }
    return true;
}

}  // namespace xdem

// xdem_calib.hpp -- one DEM step of the calibration sweep on a generic scalar S: double for
// the forward run, Dual<N> for the local Jacobians of the reverse sweep. Every function here
// evaluates on doubles exactly as the FP64 step and the reference's kernels do (the same
// operations in the same order), so the forward run of the calibration is the FP64 step, and
// on Dual<N> it returns the same values with their derivatives along the N seeded inputs.
//
// Scope (the pile of the calibration workload): Hertz-Mindlin (contact model 1) or DHS
// (model 0) without cohesion, sphere-sphere and sphere-wall contact with tangential history,
// rolling resistance, the velocity-Verlet update of position, Euler parameters and their
// rates. The parameters theta = (friction, restitution, s) enter through the material
// constants of every contact, s scaling both Young's moduli by exp(s) (s = 0: the scene's).
#pragma once

#include "xdem_dev.hpp"
#include "xdem_dual.hpp"

namespace xdem {
namespace calib {

// The material of one contact kind apart from the calibrated parameters.
struct Material {
    double Ei, Ej, pri, prj, rfric, sratio;
};

template <class S> struct Theta {
    S mu, rest, s;
};

// getConstant (impl:362-386) for a sphere pair (jr, jm > 0) or a sphere and a wall (0, 0).
template <class S> struct PairK {
    S kn, vn, ks, vs, mu;
    double eq_m, eq_r;
};

template <class S>
__host__ __device__ inline PairK<S> pair_k(int model, const Material& m, const Theta<S>& th, double ir, double jr,
                                           double im, double jm)
{
    const double pi = 3.14159265358979323846;
    PairK<S> c;
    c.eq_m = jm ? (im * jm) / (im + jm) : im;
    c.eq_r = jr ? (ir * jr) / (ir + jr) : ir;
    const S f = exp(th.s);
    const S iE = m.Ei * f, jE = m.Ej * f;
    const double ip = m.pri, jp = m.prj;
    const S Eeq = (iE * jE) / (iE * (1 - jp * jp) + jE * (1 - ip * ip));
    S Seq = (2.0 * (2.0 - ip) * (1.0 + ip) / iE) + (2.0 * (2.0 - jp) * (1.0 + jp) / jE);
    Seq = 1.0 / Seq;
    const S lne = log(th.rest);
    const S beta = -lne * sqrt(1.0 / (lne * lne + pi * pi));
    c.kn = (4.0 / 3.0) * Eeq * sqrt(c.eq_r);
    c.vn = beta;
    c.ks = model == 1 ? 8.0 * Seq : c.kn * m.sratio;
    c.vs = model == 1 ? 2.0 * sqrt(5.0 / 6.0) * beta : c.vn * m.sratio;
    c.mu = th.mu;
    return c;
}

template <class S> __host__ __device__ inline V3<S> s3(double x, double y, double z) { return {S(x), S(y), S(z)}; }
template <class S> __host__ __device__ inline V3<S> scl(double a, V3<S> v) { return {a * v.x, a * v.y, a * v.z}; }

// The force law without cohesion (impl:432-474; xContact.cpp:210-264): Fn always, Ft when the
// tangential slip direction is defined (returns true then).
template <class S>
__host__ __device__ inline bool force_law(int model, bool gpu_law, double dt, const PairK<S>& c, S cdist, S& ds,
                                          S& dots, V3<S> dv, V3<S> unit, V3<S>& Ft, V3<S>& Fn)
{
    const S fsn = -c.kn * (cdist * sqrt(cdist));
    const S dvu = dot(dv, unit);
    S fdn;
    if (model == 1 && !gpu_law)
        fdn = c.vn * (2.0 * sqrt(c.eq_m * c.kn)) * dvu;
    else
        fdn = c.vn * dvu;
    const S fnm = fsn + fdn;   // the cohesion term is -0.0 here, which adds nothing
    Fn = fnm * unit;
    const V3<S> e = dv - dvu * unit;
    const S mag_e = length(e);
    if (val(mag_e) != 0.0) {
        const V3<S> sh = e / mag_e;
        const S s_dot = dot(dv, sh);
        const S nds = ds + dt * (s_dot + dots);
        ds = nds;
        dots = s_dot;
        S ft1;
        if (model == 1) {
            const S St = c.ks * sqrt(c.eq_r * cdist);
            if (gpu_law)
                ft1 = St * nds + c.vs * sqrt(St * c.eq_m);
            else
                ft1 = St * nds + c.vs * sqrt(St * c.eq_m) * s_dot;
        } else {
            ft1 = c.ks * nds + c.vs * s_dot;
        }
        const S ft2 = c.mu * length(Fn);
        Ft = fmin(ft1, ft2) * sh;
        return true;
    }
    return false;
}

// Sums of one sphere's contacts: force, moment, rolling torque sum and rolling resistance.
template <class S> struct Sums {
    V3<S> F, M, T;
    S R;
};

template <class S> __host__ __device__ inline Sums<S> zero_sums()
{
    Sums<S> a;
    a.F = s3<S>(0, 0, 0);
    a.M = s3<S>(0, 0, 0);
    a.T = s3<S>(0, 0, 0);
    a.R = S(0.0);
    return a;
}

struct Law {
    int model;
    bool gpu_law, stale_ft;
    double dt;
};

// Sphere i's contact with sphere j (impl:613-733 for one listed pair): false (nothing added)
// when they do not overlap. sd0/sd1: the history entry in, the new one out; Ft: the
// tangential force of the previous pair of sphere i (kept when this pair's slip direction is
// undefined and the reference's stale-Ft behaviour is on).
template <class S>
__host__ __device__ inline bool pair_side(const Law& L, const PairK<S>& c, double rfric, V3<S> xi, double ir, V3<S> xj,
                                          double jr, V3<S> vi, V3<S> vj, V3<S> wi, V3<S> wj, S& sd0, S& sd1, V3<S>& Ft,
                                          Sums<S>& acc)
{
    const V3<S> rp = xj - xi;
    const S dist = length(rp);
    const S cdist = (ir + jr) - dist;
    if (!(val(cdist) > 0)) return false;
    const V3<S> unit = rp / dist;
    const V3<S> cpt = xi + S(ir) * unit;
    const V3<S> dcpr = cpt - xi;
    const V3<S> dcpr_j = cpt - xj;
    const V3<S> rv = (vj + cross(wj, dcpr_j)) - (vi + cross(wi, dcpr));
    if (!L.stale_ft) Ft = s3<S>(0, 0, 0);
    V3<S> Fn;
    force_law(L.model, L.gpu_law, L.dt, c, cdist, sd0, sd1, rv, unit, Ft, Fn);
    const V3<S> fnt = Fn + Ft;
    acc.T = acc.T + cross(dcpr, fnt);
    const double Rij = jr ? ir * jr / (ir + jr) : ir;
    acc.R = acc.R + Rij * rfric * length(Fn);
    acc.F = acc.F + fnt;
    acc.M = acc.M + cross(dcpr, fnt);
    return true;
}

// plane_detection (impl:735-815) on S: wp = (dp.u1, dp.u2, dp.uw), dp = xp - xw.
template <class S>
__host__ __device__ inline S plane_detect(const xdem_plane_info& pe, V3<S> xp, V3<S> wp, V3<S>& u, double r)
{
    const S a_l1 = (wp.x - pe.l1) * (wp.x - pe.l1);
    const S b_l2 = (wp.y - pe.l2) * (wp.y - pe.l2);
    const S sqa = wp.x * wp.x, sqb = wp.y * wp.y, sqc = wp.z * wp.z;
    const double sqr = r * r;
    const V3<S> xw = s3<S>(pe.xw.x, pe.xw.y, pe.xw.z), w2 = s3<S>(pe.w2.x, pe.w2.y, pe.w2.z);
    const V3<S> w3 = s3<S>(pe.w3.x, pe.w3.y, pe.w3.z), w4 = s3<S>(pe.w4.x, pe.w4.y, pe.w4.z);
    const double px = val(wp.x), py = val(wp.y);
    V3<S> corner;
    bool at_corner = true;
    if (px < 0 && py < 0 && val(sqa + sqb + sqc) < sqr) corner = xw;
    else if (px > pe.l1 && py < 0 && val(a_l1 + sqb + sqc) < sqr) corner = w2;
    else if (px > pe.l1 && py > pe.l2 && val(a_l1 + b_l2 + sqc) < sqr) corner = w3;
    else if (px < 0 && py > pe.l2 && val(sqa + b_l2 + sqc) < sqr) corner = w4;
    else at_corner = false;
    if (at_corner) {
        const V3<S> X = xp - corner;
        const S h = length(X);
        u = X / h;
        return r - h;
    }
    V3<S> base, tip;
    bool at_edge = true;
    if ((px > 0 && px < pe.l1) && py < 0 && val(sqb + sqc) < sqr) { base = xw; tip = w2; }
    else if ((px > 0 && px < pe.l1) && py > pe.l2 && val(b_l2 + sqc) < sqr) { base = w4; tip = w3; }
    else if ((px > 0 && py < pe.l2) && px > pe.l1 && val(a_l1 + sqc) < sqr) { base = w2; tip = w3; }
    else at_edge = false;
    if (at_edge) {
        const V3<S> X = xp - base;
        const V3<S> e = tip - base;
        const V3<S> us = e / length(e);
        const V3<S> hs = X - dot(X, us) * us;
        const S h = length(hs);
        u = (-hs) / h;
        return r - h;
    }
    if ((px > 0 && px < pe.l1) && (py > 0 && py < pe.l2)) {
        const V3<S> dp = xp - xw;
        const V3<S> uw = s3<S>(pe.uw.x, pe.uw.y, pe.uw.z);
        const V3<S> uu = uw / length(uw);
        u = -uu;
        return r - fabs(dot(dp, u));
    }
    return S(0.0);
}

// Sphere i's contact with a wall (impl:950-1080 for one sphere): false when they do not
// overlap. sd0/sd1 as for pair_side; the tangential force starts at zero for every wall.
template <class S>
__host__ __device__ inline bool plane_side(const Law& L, const PairK<S>& c, double rfric, const xdem_plane_info& pl,
                                           const double* dbi, V3<S> xi, double r, V3<S> vi, V3<S> wi, S& sd0, S& sd1,
                                           Sums<S>& acc)
{
    const V3<S> bpos = s3<S>(dbi[XDEM_PX], dbi[XDEM_PY], dbi[XDEM_PZ]);
    const V3<S> bvel = s3<S>(dbi[XDEM_VX], dbi[XDEM_VY], dbi[XDEM_VZ]);
    const double e0 = dbi[XDEM_E0], e1 = dbi[XDEM_E1], e2 = dbi[XDEM_E2], e3 = dbi[XDEM_E3];
    const double d0 = dbi[XDEM_ED0], d1 = dbi[XDEM_ED1], d2 = dbi[XDEM_ED2], d3 = dbi[XDEM_ED3];
    const V3<S> oj = s3<S>(2.0 * (-e1 * d0 + e0 * d1 - e3 * d2 + e2 * d3), 2.0 * (-e2 * d0 + e3 * d1 + e0 * d2 - e1 * d3),
                           2.0 * (-e3 * d0 - e2 * d1 + e1 * d2 + e0 * d3));
    const V3<S> dp = xi - s3<S>(pl.xw.x, pl.xw.y, pl.xw.z);
    const V3<S> wp = {dot(dp, s3<S>(pl.u1.x, pl.u1.y, pl.u1.z)), dot(dp, s3<S>(pl.u2.x, pl.u2.y, pl.u2.z)),
                      dot(dp, s3<S>(pl.uw.x, pl.uw.y, pl.uw.z))};
    V3<S> unit = s3<S>(0, 0, 0);
    const S cdist = plane_detect(pl, xi, wp, unit, r);
    if (!(val(cdist) > 0)) return false;
    const V3<S> cpt = xi + S(r) * unit;
    const V3<S> dcpr_j = cpt - bpos;
    const V3<S> dcpr = cpt - xi;
    const V3<S> dv = (bvel + cross(oj, dcpr_j)) - (vi + cross(wi, dcpr));
    V3<S> Ft = s3<S>(0, 0, 0), Fn;
    force_law(L.model, L.gpu_law, L.dt, c, cdist, sd0, sd1, dv, unit, Ft, Fn);
    const V3<S> mf = Fn + Ft;
    acc.T = acc.T + cross(dcpr, mf);
    acc.R = acc.R + r * rfric * length(Fn);
    acc.F = acc.F + mf;
    acc.M = acc.M + cross(dcpr, mf);
    return true;
}

// omega = 2 G(e) e_dot (impl:34-42)
template <class S> __host__ __device__ inline V3<S> ang_vel(V4<S> e, V4<S> d)
{
    const V3<S> o = {-e.y * d.x + e.x * d.y - e.w * d.z + e.z * d.w, -e.z * d.x + e.w * d.y + e.x * d.z - e.y * d.w,
                     -e.w * d.x - e.z * d.y + e.y * d.z + e.x * d.w};
    return scl(2.0, o);
}

// Rolling resistance (impl:2420-2454): adds the resisting moment to M when Tr != 0.
template <class S>
__host__ __device__ inline void rolling(double dt, double J, V3<S> w, V3<S> Tm, S Tr, V3<S>& M)
{
    if (!(val(Tr) != 0.0)) return;
    const double Jdt = J * dt;
    const V3<S> T = {Jdt * w.x - Tm.x, Jdt * w.y - Tm.y, Jdt * w.z - Tm.z};
    const S lT = length(T);
    if (val(lT) != 0.0) {
        V3<S> Trv = Tr * (T / lT);
        if (val(length(Trv)) >= val(length(T))) Trv = T;
        M = M + Trv;
    }
}

// A(e)^T v (impl:8-52)
template <class S> __host__ __device__ inline V3<S> to_local(V3<S> v, V4<S> ep)
{
    const V3<S> r0 = {2.0 * (ep.x * ep.x + ep.y * ep.y - 0.5), 2.0 * (ep.y * ep.z - ep.x * ep.w), 2.0 * (ep.y * ep.w + ep.x * ep.z)};
    const V3<S> r1 = {2.0 * (ep.y * ep.z + ep.x * ep.w), 2.0 * (ep.x * ep.x + ep.z * ep.z - 0.5), 2.0 * (ep.z * ep.w - ep.x * ep.y)};
    const V3<S> r2 = {2.0 * (ep.y * ep.w - ep.x * ep.z), 2.0 * (ep.z * ep.w + ep.x * ep.y), 2.0 * (ep.x * ep.x + ep.w * ep.w - 0.5)};
    return {r0.x * v.x + r1.x * v.y + r2.x * v.z, r0.y * v.x + r1.y * v.y + r2.y * v.z, r0.z * v.x + r1.z * v.y + r2.z * v.z};
}

// calculate_uceom (impl:74-133): the Euler parameter accelerations from the body-frame moment.
template <class S> __host__ __device__ inline V4<S> uceom(V3<double> J, V4<S> ep, V4<S> ev, V3<S> n)
{
    S a[4][4];
    a[0][0] = (2.0 * J.x) * (-ep.y); a[0][1] = (2.0 * J.x) * ep.x;    a[0][2] = (2.0 * J.x) * ep.w;    a[0][3] = (2.0 * J.x) * (-ep.z);
    a[1][0] = (2.0 * J.y) * (-ep.z); a[1][1] = (2.0 * J.y) * (-ep.w); a[1][2] = (2.0 * J.y) * ep.x;    a[1][3] = (2.0 * J.y) * ep.y;
    a[2][0] = (2.0 * J.z) * (-ep.w); a[2][1] = (2.0 * J.z) * ep.z;    a[2][2] = (2.0 * J.z) * (-ep.y); a[2][3] = (2.0 * J.z) * ep.x;
    a[3][0] = ep.x;                  a[3][1] = ep.y;                  a[3][2] = ep.z;                  a[3][3] = ep.w;
    const V3<S> T = {J.x * (-ep.y * ev.x + ep.x * ev.y + ep.w * ev.z - ep.z * ev.w),
                     J.y * (-ep.z * ev.x - ep.w * ev.y + ep.x * ev.z + ep.y * ev.w),
                     J.z * (-ep.w * ev.x + ep.z * ev.y - ep.y * ev.z + ep.x * ev.w)};
    V4<S> L0;
    L0.x = 4.0 * (-ev.y * T.x - ev.z * T.y - ev.w * T.z);
    L0.y = 4.0 * (ev.x * T.x - ev.w * T.y + ev.z * T.z);
    L0.z = 4.0 * (ev.w * T.x + ev.x * T.y - ev.y * T.z);
    L0.w = 4.0 * (-ev.z * T.x + ev.y * T.y + ev.x * T.z);
    const V3<S> LH = {-ep.y * L0.x + ep.x * L0.y + ep.w * L0.z - ep.z * L0.w,
                      -ep.z * L0.x - ep.w * L0.y + ep.x * L0.z + ep.y * L0.w,
                      -ep.w * L0.x + ep.z * L0.y - ep.y * L0.z + ep.x * L0.w};
    const S evev = ev.x * ev.x + ev.y * ev.y + ev.z * ev.z + ev.w * ev.w;
    const S rhs[4] = {n.x - LH.x, n.y - LH.y, n.z - LH.z, 0.0 - evev};
#define A_(i, j) a[i][j]
    const S det =
        A_(0,0)*A_(1,1)*A_(2,2)*A_(3,3) + A_(0,0)*A_(1,2)*A_(2,3)*A_(3,1) + A_(0,0)*A_(1,3)*A_(2,1)*A_(3,2) -
        A_(0,0)*A_(1,3)*A_(2,2)*A_(3,1) - A_(0,0)*A_(1,2)*A_(2,1)*A_(3,3) - A_(0,0)*A_(1,1)*A_(2,3)*A_(3,2) -
        A_(0,1)*A_(1,0)*A_(2,2)*A_(3,3) - A_(0,2)*A_(1,0)*A_(2,3)*A_(3,1) - A_(0,3)*A_(1,0)*A_(2,1)*A_(3,2) +
        A_(0,3)*A_(1,0)*A_(2,2)*A_(3,1) + A_(0,2)*A_(1,0)*A_(2,1)*A_(3,3) + A_(0,1)*A_(1,0)*A_(2,3)*A_(3,2) +
        A_(0,1)*A_(1,2)*A_(2,0)*A_(3,3) + A_(0,2)*A_(1,3)*A_(2,0)*A_(3,1) + A_(0,3)*A_(1,1)*A_(2,0)*A_(3,2) -
        A_(0,3)*A_(1,2)*A_(2,0)*A_(3,1) - A_(0,2)*A_(1,1)*A_(2,0)*A_(3,3) - A_(0,1)*A_(1,3)*A_(2,0)*A_(3,2) -
        A_(0,1)*A_(1,2)*A_(2,3)*A_(3,0) - A_(0,2)*A_(1,3)*A_(2,1)*A_(3,0) - A_(0,3)*A_(1,1)*A_(2,2)*A_(3,0) +
        A_(0,3)*A_(1,2)*A_(2,1)*A_(3,0) + A_(0,2)*A_(1,1)*A_(2,3)*A_(3,0) + A_(0,1)*A_(1,3)*A_(2,2)*A_(3,0);
    S o[4][4];
    o[0][0] =  A_(1,1)*A_(2,2)*A_(3,3) + A_(1,2)*A_(2,3)*A_(3,1) + A_(1,3)*A_(2,1)*A_(3,2) - A_(1,3)*A_(2,2)*A_(3,1) - A_(1,2)*A_(2,1)*A_(3,3) - A_(1,1)*A_(2,3)*A_(3,2);
    o[0][1] = -A_(0,1)*A_(2,2)*A_(3,3) - A_(0,2)*A_(2,3)*A_(3,1) - A_(0,3)*A_(2,1)*A_(3,2) + A_(0,3)*A_(2,2)*A_(3,1) + A_(0,2)*A_(2,1)*A_(3,3) + A_(0,1)*A_(2,3)*A_(3,2);
    o[0][2] =  A_(0,1)*A_(1,2)*A_(3,3) + A_(0,2)*A_(1,3)*A_(3,1) + A_(0,3)*A_(1,1)*A_(3,2) - A_(0,3)*A_(1,2)*A_(3,1) - A_(0,2)*A_(1,1)*A_(3,3) - A_(0,1)*A_(1,3)*A_(3,2);
    o[0][3] = -A_(0,1)*A_(1,2)*A_(2,3) - A_(0,2)*A_(1,3)*A_(2,1) - A_(0,3)*A_(1,1)*A_(2,2) + A_(0,3)*A_(1,2)*A_(2,1) + A_(0,2)*A_(1,1)*A_(2,3) + A_(0,1)*A_(1,3)*A_(2,2);
    o[1][0] = -A_(1,0)*A_(2,2)*A_(3,3) - A_(1,2)*A_(2,3)*A_(3,0) - A_(1,3)*A_(2,0)*A_(3,2) + A_(1,3)*A_(2,2)*A_(3,0) + A_(1,2)*A_(2,0)*A_(3,3) + A_(1,0)*A_(2,3)*A_(3,2);
    o[1][1] =  A_(0,0)*A_(2,2)*A_(3,3) + A_(0,2)*A_(2,3)*A_(3,0) + A_(0,3)*A_(2,0)*A_(3,2) - A_(0,3)*A_(2,2)*A_(3,0) - A_(0,2)*A_(2,0)*A_(3,3) - A_(0,0)*A_(2,3)*A_(3,2);
    o[1][2] = -A_(0,0)*A_(1,2)*A_(3,3) - A_(0,2)*A_(1,3)*A_(3,0) - A_(0,3)*A_(1,0)*A_(3,2) + A_(0,3)*A_(1,2)*A_(3,0) + A_(0,2)*A_(1,0)*A_(3,3) + A_(0,0)*A_(1,3)*A_(3,2);
    o[1][3] =  A_(0,0)*A_(1,2)*A_(2,3) + A_(0,2)*A_(1,3)*A_(2,0) + A_(0,3)*A_(1,0)*A_(2,2) - A_(0,3)*A_(1,2)*A_(2,0) - A_(0,2)*A_(1,0)*A_(2,3) - A_(0,0)*A_(1,3)*A_(2,2);
    o[2][0] =  A_(1,0)*A_(2,1)*A_(3,3) + A_(1,1)*A_(2,3)*A_(3,0) + A_(1,3)*A_(2,0)*A_(3,1) - A_(1,3)*A_(2,1)*A_(3,0) - A_(1,1)*A_(2,0)*A_(3,3) - A_(1,0)*A_(2,3)*A_(3,1);
    o[2][1] = -A_(0,0)*A_(2,1)*A_(3,3) - A_(0,1)*A_(2,3)*A_(3,0) - A_(0,3)*A_(2,0)*A_(3,1) + A_(0,3)*A_(2,1)*A_(3,0) + A_(0,1)*A_(2,0)*A_(3,3) + A_(0,0)*A_(2,3)*A_(3,1);
    o[2][2] =  A_(0,0)*A_(1,1)*A_(3,3) + A_(0,1)*A_(1,3)*A_(3,0) + A_(0,3)*A_(1,0)*A_(3,1) - A_(0,3)*A_(1,1)*A_(3,0) - A_(0,1)*A_(1,0)*A_(3,3) - A_(0,0)*A_(1,3)*A_(3,1);
    o[2][3] = -A_(0,0)*A_(1,1)*A_(2,3) - A_(0,1)*A_(1,3)*A_(2,0) - A_(0,3)*A_(1,0)*A_(2,1) + A_(0,3)*A_(1,1)*A_(2,0) + A_(0,1)*A_(1,0)*A_(2,3) + A_(0,0)*A_(1,3)*A_(2,1);
    o[3][0] = -A_(1,0)*A_(2,1)*A_(3,2) - A_(1,1)*A_(2,2)*A_(3,0) - A_(1,2)*A_(2,0)*A_(3,1) + A_(1,2)*A_(2,1)*A_(3,0) + A_(1,1)*A_(2,0)*A_(3,2) + A_(1,0)*A_(2,2)*A_(3,1);
    o[3][1] =  A_(0,0)*A_(2,1)*A_(3,2) + A_(0,1)*A_(2,2)*A_(3,0) + A_(0,2)*A_(2,0)*A_(3,1) - A_(0,2)*A_(2,1)*A_(3,0) - A_(0,1)*A_(2,0)*A_(3,2) - A_(0,0)*A_(2,2)*A_(3,1);
    o[3][2] = -A_(0,0)*A_(1,1)*A_(3,2) - A_(0,1)*A_(1,2)*A_(3,0) - A_(0,2)*A_(1,0)*A_(3,1) + A_(0,2)*A_(1,1)*A_(3,0) + A_(0,1)*A_(1,0)*A_(3,2) + A_(0,0)*A_(1,2)*A_(3,1);
    o[3][3] =  A_(0,0)*A_(1,1)*A_(2,2) + A_(0,1)*A_(1,2)*A_(2,0) + A_(0,2)*A_(1,0)*A_(2,1) - A_(0,2)*A_(1,1)*A_(2,0) - A_(0,1)*A_(1,0)*A_(2,2) - A_(0,0)*A_(1,2)*A_(2,1);
#undef A_
    const S m = 1.0 / det;
    V4<S> r;
    r.x = m * (o[0][0] * rhs[0] + o[0][1] * rhs[1] + o[0][2] * rhs[2] + o[0][3] * rhs[3]);
    r.y = m * (o[1][0] * rhs[0] + o[1][1] * rhs[1] + o[1][2] * rhs[2] + o[1][3] * rhs[3]);
    r.z = m * (o[2][0] * rhs[0] + o[2][1] * rhs[1] + o[2][2] * rhs[2] + o[2][3] * rhs[3]);
    r.w = m * (o[3][0] * rhs[0] + o[3][1] * rhs[1] + o[3][2] * rhs[2] + o[3][3] * rhs[3]);
    return r;
}

// The velocity update of one sphere (xIntegratorVV.cpp:131-152; impl:192-222) from its
// contact sums (rolling resistance already in M): v, acc, ev, ea updated in place.
template <class S>
__host__ __device__ inline void velocity_update(double dt, V3<double> g, double m, V3<double> J, V4<S> e, V3<S> F,
                                                V3<S> M, V3<S>& v, V3<S>& acc, V4<S>& ev, V4<S>& ea)
{
    const V3<S> np_ = to_local(M, e);
    const V4<S> mea = uceom(J, e, ev, np_);
    const double inv_m = 1.0 / m;
    const V3<S> a = {inv_m * (F.x + m * g.x), inv_m * (F.y + m * g.y), inv_m * (F.z + m * g.z)};
    const double hdt = 0.5 * dt;
    v = {v.x + hdt * acc.x, v.y + hdt * acc.y, v.z + hdt * acc.z};
    ev = {ev.x + hdt * ea.x, ev.y + hdt * ea.y, ev.z + hdt * ea.z, ev.w + hdt * ea.w};
    v = {v.x + hdt * a.x, v.y + hdt * a.y, v.z + hdt * a.z};
    ev = {ev.x + hdt * mea.x, ev.y + hdt * mea.y, ev.z + hdt * mea.z, ev.w + hdt * mea.w};
    acc = a;
    ea = mea;
}

// The position update of one sphere (impl:135-190): position and normalised Euler parameters.
template <class S>
__host__ __device__ inline void position_update(double dt, double h, V3<S>& p, V4<S>& e, V3<S> v, V4<S> d, V3<S> ac,
                                                V4<S> a)
{
    V4<S> ne = {e.x + dt * d.x + h * a.x, e.y + dt * d.y + h * a.y, e.z + dt * d.z + h * a.z, e.w + dt * d.w + h * a.w};
    const S l = sqrt(ne.x * ne.x + ne.y * ne.y + ne.z * ne.z + ne.w * ne.w);
    e = {ne.x / l, ne.y / l, ne.z / l, ne.w / l};
    const V3<S> dp = {dt * v.x + h * ac.x, dt * v.y + h * ac.y, dt * v.z + h * ac.z};
    p = {p.x + dp.x, p.y + dp.y, p.z + dp.z};
}

}  // namespace calib
}  // namespace xdem

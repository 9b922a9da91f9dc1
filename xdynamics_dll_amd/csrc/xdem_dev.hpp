// xdem_dev.hpp -- device-side math and contact physics of the DEM step (gfx950).
//
// Templated on the arithmetic type T (double for the FP64 path, float for the mixed path).
// Expression order follows the reference kernel text
// (include/xdynamics_parallel/xParallelDEM_impl.cuh, "impl" below) and, for the force law
// the build uses by default, source/xdynamics_object/xContact.cpp:210-264, so that the FP64
// path computed with -ffp-contract=off rounds exactly like a plain C restatement.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "xdem_types.h"

namespace xdem {

template <class T> struct V3 { T x, y, z; };
template <class T> struct V4 { T x, y, z, w; };

template <class T> __host__ __device__ __forceinline__ V3<T> mk3(T x, T y, T z) { return {x, y, z}; }
template <class T> __host__ __device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __host__ __device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __host__ __device__ __forceinline__ V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> __host__ __device__ __forceinline__ V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> __host__ __device__ __forceinline__ V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> __host__ __device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __host__ __device__ __forceinline__ T length(V3<T> a) { return sqrt(dot(a, a)); }
template <class T> __host__ __device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T, class U> __host__ __device__ __forceinline__ V3<T> cvt3(V3<U> a) { return {(T)a.x, (T)a.y, (T)a.z}; }

__device__ __forceinline__ V3<double> ld3(const double* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ V4<double> ld4(const double* p)
{
    const double2 a = reinterpret_cast<const double2*>(p)[0];
    const double2 b = reinterpret_cast<const double2*>(p)[1];
    return {a.x, a.y, b.x, b.y};
}
__device__ __forceinline__ void st3(double* p, V3<double> a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
__device__ __forceinline__ void st4(double* p, V4<double> a)
{
    reinterpret_cast<double2*>(p)[0] = make_double2(a.x, a.y);
    reinterpret_cast<double2*>(p)[1] = make_double2(a.z, a.w);
}

// ------------------------------------------------------------------ rotations (impl:8-52)
__device__ __forceinline__ V3<double> rot_global(V3<double> v, V4<double> ep)
{
    V3<double> r0 = mk3(2.0 * (ep.x * ep.x + ep.y * ep.y - 0.5), 2.0 * (ep.y * ep.z - ep.x * ep.w), 2.0 * (ep.y * ep.w + ep.x * ep.z));
    V3<double> r1 = mk3(2.0 * (ep.y * ep.z + ep.x * ep.w), 2.0 * (ep.x * ep.x + ep.z * ep.z - 0.5), 2.0 * (ep.z * ep.w - ep.x * ep.y));
    V3<double> r2 = mk3(2.0 * (ep.y * ep.w - ep.x * ep.z), 2.0 * (ep.z * ep.w + ep.x * ep.y), 2.0 * (ep.x * ep.x + ep.w * ep.w - 0.5));
    return mk3(r0.x * v.x + r0.y * v.y + r0.z * v.z, r1.x * v.x + r1.y * v.y + r1.z * v.z, r2.x * v.x + r2.y * v.y + r2.z * v.z);
}

__device__ __forceinline__ V3<double> rot_local(V3<double> v, V4<double> ep)
{
    V3<double> r0 = mk3(2.0 * (ep.x * ep.x + ep.y * ep.y - 0.5), 2.0 * (ep.y * ep.z - ep.x * ep.w), 2.0 * (ep.y * ep.w + ep.x * ep.z));
    V3<double> r1 = mk3(2.0 * (ep.y * ep.z + ep.x * ep.w), 2.0 * (ep.x * ep.x + ep.z * ep.z - 0.5), 2.0 * (ep.z * ep.w - ep.x * ep.y));
    V3<double> r2 = mk3(2.0 * (ep.y * ep.w - ep.x * ep.z), 2.0 * (ep.z * ep.w + ep.x * ep.y), 2.0 * (ep.x * ep.x + ep.w * ep.w - 0.5));
    return mk3(r0.x * v.x + r1.x * v.y + r2.x * v.z, r0.y * v.x + r1.y * v.y + r2.y * v.z, r0.z * v.x + r1.z * v.y + r2.z * v.z);
}

// omega = 2 L(e) e_dot
__device__ __forceinline__ V3<double> omega_of(V4<double> e, V4<double> d)
{
    V3<double> o = mk3(-e.y * d.x + e.x * d.y - e.w * d.z + e.z * d.w,
                       -e.z * d.x + e.w * d.y + e.x * d.z - e.y * d.w,
                       -e.w * d.x - e.z * d.y + e.y * d.z + e.x * d.w);
    return 2.0 * o;
}

// ------------------------------------------------------ constrained rotational EOM (impl:74-133)
// Solves [2 J G(e); e^T] e_dd = [n' - G(e) 4 G(e_d)^T J G(e) e_d ; -e_d.e_d] by the adjugate
// of the 4x4 matrix; the determinant and cofactor sums keep the reference's term order.
__device__ __forceinline__ V4<double> uceom(V3<double> J, V4<double> ep, V4<double> ev, V3<double> n)
{
    const double a00 = 2.0 * J.x * (-ep.y), a01 = 2.0 * J.x * ep.x, a02 = 2.0 * J.x * ep.w, a03 = 2.0 * J.x * (-ep.z);
    const double a10 = 2.0 * J.y * (-ep.z), a11 = 2.0 * J.y * (-ep.w), a12 = 2.0 * J.y * ep.x, a13 = 2.0 * J.y * ep.y;
    const double a20 = 2.0 * J.z * (-ep.w), a21 = 2.0 * J.z * ep.z, a22 = 2.0 * J.z * (-ep.y), a23 = 2.0 * J.z * ep.x;
    const double a30 = ep.x, a31 = ep.y, a32 = ep.z, a33 = ep.w;
    const V3<double> T = mk3(J.x * (-ep.y * ev.x + ep.x * ev.y + ep.w * ev.z - ep.z * ev.w),
                             J.y * (-ep.z * ev.x - ep.w * ev.y + ep.x * ev.z + ep.y * ev.w),
                             J.z * (-ep.w * ev.x + ep.z * ev.y - ep.y * ev.z + ep.x * ev.w));
    const double l0 = 4.0 * (-ev.y * T.x - ev.z * T.y - ev.w * T.z);
    const double l1 = 4.0 * (ev.x * T.x - ev.w * T.y + ev.z * T.z);
    const double l2 = 4.0 * (ev.w * T.x + ev.x * T.y - ev.y * T.z);
    const double l3 = 4.0 * (-ev.z * T.x + ev.y * T.y + ev.x * T.z);
    const double h0 = -ep.y * l0 + ep.x * l1 + ep.w * l2 - ep.z * l3;
    const double h1 = -ep.z * l0 - ep.w * l1 + ep.x * l2 + ep.y * l3;
    const double h2 = -ep.w * l0 + ep.z * l1 - ep.y * l2 + ep.x * l3;
    const double ee = ev.x * ev.x + ev.y * ev.y + ev.z * ev.z + ev.w * ev.w;
    const double r0 = n.x - h0, r1 = n.y - h1, r2 = n.z - h2, r3 = 0.0 - ee;
    const double det =
        a00*a11*a22*a33 + a00*a12*a23*a31 + a00*a13*a21*a32 -
        a00*a13*a22*a31 - a00*a12*a21*a33 - a00*a11*a23*a32 -
        a01*a10*a22*a33 - a02*a10*a23*a31 - a03*a10*a21*a32 +
        a03*a10*a22*a31 + a02*a10*a21*a33 + a01*a10*a23*a32 +
        a01*a12*a20*a33 + a02*a13*a20*a31 + a03*a11*a20*a32 -
        a03*a12*a20*a31 - a02*a11*a20*a33 - a01*a13*a20*a32 -
        a01*a12*a23*a30 - a02*a13*a21*a30 - a03*a11*a22*a30 +
        a03*a12*a21*a30 + a02*a11*a23*a30 + a01*a13*a22*a30;
    const double o00 =  a11*a22*a33 + a12*a23*a31 + a13*a21*a32 - a13*a22*a31 - a12*a21*a33 - a11*a23*a32;
    const double o01 = -a01*a22*a33 - a02*a23*a31 - a03*a21*a32 + a03*a22*a31 + a02*a21*a33 + a01*a23*a32;
    const double o02 =  a01*a12*a33 + a02*a13*a31 + a03*a11*a32 - a03*a12*a31 - a02*a11*a33 - a01*a13*a32;
    const double o03 = -a01*a12*a23 - a02*a13*a21 - a03*a11*a22 + a03*a12*a21 + a02*a11*a23 + a01*a13*a22;
    const double o10 = -a10*a22*a33 - a12*a23*a30 - a13*a20*a32 + a13*a22*a30 + a12*a20*a33 + a10*a23*a32;
    const double o11 =  a00*a22*a33 + a02*a23*a30 + a03*a20*a32 - a03*a22*a30 - a02*a20*a33 - a00*a23*a32;
    const double o12 = -a00*a12*a33 - a02*a13*a30 - a03*a10*a32 + a03*a12*a30 + a02*a10*a33 + a00*a13*a32;
    const double o13 =  a00*a12*a23 + a02*a13*a20 + a03*a10*a22 - a03*a12*a20 - a02*a10*a23 - a00*a13*a22;
    const double o20 =  a10*a21*a33 + a11*a23*a30 + a13*a20*a31 - a13*a21*a30 - a11*a20*a33 - a10*a23*a31;
    const double o21 = -a00*a21*a33 - a01*a23*a30 - a03*a20*a31 + a03*a21*a30 + a01*a20*a33 + a00*a23*a31;
    const double o22 =  a00*a11*a33 + a01*a13*a30 + a03*a10*a31 - a03*a11*a30 - a01*a10*a33 - a00*a13*a31;
    const double o23 = -a00*a11*a23 - a01*a13*a20 - a03*a10*a21 + a03*a11*a20 + a01*a10*a23 + a00*a13*a21;
    const double o30 = -a10*a21*a32 - a11*a22*a30 - a12*a20*a31 + a12*a21*a30 + a11*a20*a32 + a10*a22*a31;
    const double o31 =  a00*a21*a32 + a01*a22*a30 + a02*a20*a31 - a02*a21*a30 - a01*a20*a32 - a00*a22*a31;
    const double o32 = -a00*a11*a32 - a01*a12*a30 - a02*a10*a31 + a02*a11*a30 + a01*a10*a32 + a00*a12*a31;
    const double o33 =  a00*a11*a22 + a01*a12*a20 + a02*a10*a21 - a02*a11*a20 - a01*a10*a22 - a00*a12*a21;
    const double m = 1.0 / det;
    V4<double> r;
    r.x = m * (o00 * r0 + o01 * r1 + o02 * r2 + o03 * r3);
    r.y = m * (o10 * r0 + o11 * r1 + o12 * r2 + o13 * r3);
    r.z = m * (o20 * r0 + o21 * r1 + o22 * r2 + o23 * r3);
    r.w = m * (o30 * r0 + o31 * r1 + o32 * r2 + o33 * r3);
    return r;
}

// ----------------------------------------------------------------- contact constants
// Material-pair constants of getConstant (impl:362-386) and the cohesion helpers
// (impl:388-430) that do not depend on the two radii or masses; the host fills them
// once per contact object, with the same operation order as the per-pair formulas.
struct ContactConst {
    double k43E;    // (4/3) E*            -> kn = k43E * sqrt(R*)
    double beta;    // -ln e / sqrt(ln^2 e + pi^2)  (vn)
    double ks_hm;   // 8 G*
    double vs_hm;   // 2 sqrt(5/6) beta
    double vs_dhs;  // beta * stiffness ratio
    double sratio;
    double fric;
    double rfric;
    double coh;
    double c3cp;    // (3 coh) pi           -> c0 = c3cp * R*
    double E4;      // 4 E*_c
    double c8;      // ((8 pi) coh) E*_c
    double pcc;     // ((pi pi) coh) coh
    double ee;      // E*_c E*_c
    double jk;      // (-(3/2) pi) coh      -> JKR pull-off = jk * R*
    double coh_reach; // upper bound of |cohesion separation depth| for this pair (0 without cohesion)
};

// Per-pair constants at precision T.
template <class T> struct PairConst {
    T kn, vn, ks, vs, mu, eq_r, eq_m;
    double eq_r_d, eq_m_d;  // effective radius and mass in FP64
};

// Cohesion force (impl:388-413). Req is the pair's effective radius, Fn the normal force.
template <class T>
__device__ __forceinline__ T cohesion_force(const ContactConst& cc, T Req, T Fn)
{
    T cf = T(0);
    if (cc.coh) {
        const T c0 = T(cc.c3cp) * Req;
        const T eq = T(2.0) * c0 * Fn + c0 * c0;
        if (eq <= T(0)) Fn = T(-0.5) * c0;
        const T a3 = (T(3.0) * Req) * (Fn + c0 + sqrt(fabs(T(2.0) * c0 * Fn + c0 * c0))) / T(cc.E4);
        cf = -sqrt(T(cc.c8) * a3);
    }
    return cf;
}

// Separation depth below which a cohesive pair still attracts (impl:415-423).
__device__ __forceinline__ double cohesion_limit_depth(const ContactConst& cc, double Req)
{
    const double c1 = (cc.pcc * Req) / cc.ee;
    return -(3.0 / 4.0) * pow(c1, 1.0 / 3.0);
}

// Normal and tangential force of one pair (impl:432-474; xContact.cpp:210-264).
//   MODEL 1: Hertz-Mindlin no slip, MODEL 0: DHS.
//   GPU_LAW selects the kernel's damping terms (impl) instead of xContact.cpp's.
// ds/dots is the tangential history, kept in FP64 for every T. Ft is updated only
// when the tangential slip direction is defined (|e| != 0), as in the reference.
//
// The relative velocity dv, the contact normal, the normal force (Hertz + damping + JKR)
// and the split of dv into normal and tangential parts are FP64 for every T: near contact
// onset the damping and cohesive terms cancel to ~1e-3 of either, and when dv is nearly
// normal the tangential slip is a difference of nearly equal velocities, so FP32 there
// would miss the 1e-4 force bound by orders of magnitude. With T = float (the mixed path)
// the tangential spring-damper magnitude, its Coulomb cap and the output vectors are FP32.
// Returns true when Ft was set (|e| != 0).
template <class T, int MODEL, bool GPU_LAW>
__device__ __forceinline__ bool force_law(const ContactConst& cc, const PairConst<T>& c, double cdist, double dt,
                                          double& ds, double& dots, V3<double> dv, V3<double> unit, V3<T>& Ft,
                                          V3<T>& Fn)
{
    const double Req = c.eq_r_d, Meq = c.eq_m_d;
    const double kn = std::is_same<T, double>::value ? (double)c.kn : cc.k43E * sqrt(Req);
    const double fsn = -kn * (cdist * sqrt(cdist));
    const double dvu = dot(dv, unit);
    double fdn;
    if (MODEL == 1 && !GPU_LAW)
        fdn = cc.beta * (2.0 * sqrt(Meq * kn)) * dvu;
    else
        fdn = cc.beta * dvu;
    const double fca = -cohesion_force<double>(cc, Req, fsn + fdn);
    const double fnm = !GPU_LAW ? (fsn + fdn + fca) : (fsn + fca + fdn);
    const V3<double> Fn_d = fnm * unit;
    Fn = cvt3<T>(Fn_d);
    const V3<double> e = dv - dvu * unit;
    const double mag_e = length(e);
    if (mag_e) {
        const V3<double> sh = e / mag_e;
        const double s_dot = dot(dv, sh);
        const double nds = ds + dt * (s_dot + dots);
        ds = nds;
        dots = s_dot;
        T ft1;
        if (MODEL == 1) {
            const T St = c.ks * sqrt(c.eq_r * T(cdist));
            if (GPU_LAW)
                ft1 = St * T(nds) + c.vs * sqrt(St * c.eq_m);
            else
                ft1 = St * T(nds) + c.vs * sqrt(St * c.eq_m) * T(s_dot);
        } else {
            ft1 = c.ks * T(nds) + c.vs * T(s_dot);
        }
        const T ft2 = c.mu * length(Fn);
        Ft = fmin(ft1, ft2) * cvt3<T>(sh);
        return true;
    }
    return false;
}

template <class T, int MODEL>
__device__ __forceinline__ PairConst<T> pair_const(const ContactConst& cc, double ir, double jr, double im, double jm)
{
    PairConst<T> c;
    const double Meq = jm ? (im * jm) / (im + jm) : im;
    const double Req = jr ? (ir * jr) / (ir + jr) : ir;
    const T Req_t = T(Req);
    c.eq_m = T(Meq);
    c.eq_r = Req_t;
    c.eq_m_d = Meq;
    c.eq_r_d = Req;
    c.kn = T(cc.k43E) * sqrt(Req_t);
    c.vn = T(cc.beta);
    c.ks = MODEL == 1 ? T(cc.ks_hm) : c.kn * T(cc.sratio);
    c.vs = MODEL == 1 ? T(cc.vs_hm) : T(cc.vs_dhs);
    c.mu = T(cc.fric);
    return c;
}

// ------------------------------------------------------------- wall detection (impl:735-815)
// wp = (dp.u1, dp.u2, dp.uw) with dp = xp - xw. Returns the overlap and sets u; returns 0
// (u untouched) when no region matches. Corner regions set u from the corner to the
// sphere, edges and the face from the sphere towards the wall, as the kernel text does.
__device__ __forceinline__ double plane_detection(const xdem_plane_info& pe, V3<double> xp, V3<double> wp,
                                                  V3<double>& u, double r)
{
    const double a_l1 = (wp.x - pe.l1) * (wp.x - pe.l1);
    const double b_l2 = (wp.y - pe.l2) * (wp.y - pe.l2);
    const double sqa = wp.x * wp.x, sqb = wp.y * wp.y, sqc = wp.z * wp.z, sqr = r * r;
    const V3<double> xw = mk3(pe.xw.x, pe.xw.y, pe.xw.z), w2 = mk3(pe.w2.x, pe.w2.y, pe.w2.z);
    const V3<double> w3 = mk3(pe.w3.x, pe.w3.y, pe.w3.z), w4 = mk3(pe.w4.x, pe.w4.y, pe.w4.z);
    V3<double> corner;
    bool at_corner = true;
    if (wp.x < 0 && wp.y < 0 && (sqa + sqb + sqc) < sqr) corner = xw;
    else if (wp.x > pe.l1 && wp.y < 0 && (a_l1 + sqb + sqc) < sqr) corner = w2;
    else if (wp.x > pe.l1 && wp.y > pe.l2 && (a_l1 + b_l2 + sqc) < sqr) corner = w3;
    else if (wp.x < 0 && wp.y > pe.l2 && (sqa + b_l2 + sqc) < sqr) corner = w4;
    else at_corner = false;
    if (at_corner) {
        const V3<double> X = xp - corner;
        const double h = length(X);
        u = X / h;
        return r - h;
    }
    V3<double> base, tip;
    bool at_edge = true;
    if ((wp.x > 0 && wp.x < pe.l1) && wp.y < 0 && (sqb + sqc) < sqr) { base = xw; tip = w2; }
    else if ((wp.x > 0 && wp.x < pe.l1) && wp.y > pe.l2 && (b_l2 + sqc) < sqr) { base = w4; tip = w3; }
    else if ((wp.x > 0 && wp.y < pe.l2) && wp.x > pe.l1 && (a_l1 + sqc) < sqr) { base = w2; tip = w3; }
    else at_edge = false;
    if (at_edge) {
        const V3<double> X = xp - base;
        const V3<double> e = tip - base;
        const V3<double> us = e / length(e);
        const V3<double> hs = X - dot(X, us) * us;
        const double h = length(hs);
        u = (-hs) / h;
        return r - h;
    }
    if ((wp.x > 0 && wp.x < pe.l1) && (wp.y > 0 && wp.y < pe.l2)) {
        const V3<double> dp = xp - xw;
        const V3<double> uw = mk3(pe.uw.x, pe.uw.y, pe.uw.z);
        const V3<double> uu = uw / length(uw);
        u = -uu;
        return r - fabs(dot(dp, u));
    }
    return 0.0;
}

// ----------------------------------------------------- closest point on a triangle (impl:1655-1705)
// All region tests run in order and a later match overrides an earlier one (vertex,
// then edge, then face); ct = 0 vertex, 1 edge, 2 face, -1 none.
__device__ __forceinline__ V3<double> closest_on_triangle(const xdem_triangle_info& t, V3<double> p, int& ct)
{
    const V3<double> a = mk3(t.P.x, t.P.y, t.P.z), b = mk3(t.Q.x, t.Q.y, t.Q.z), c = mk3(t.R.x, t.R.y, t.R.z);
    const V3<double> ab = b - a, ac = c - a, ap = p - a, bp = p - b, cp = p - c;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap), d3 = dot(ab, bp);
    const double d4 = dot(ac, bp), d5 = dot(ab, cp), d6 = dot(ac, cp);
    const double va = d3 * d6 - d5 * d4;
    const double vb = d5 * d2 - d1 * d6;
    const double vc = d1 * d4 - d3 * d2;
    V3<double> q = mk3(0.0, 0.0, 0.0);
    ct = -1;
    if (d1 <= 0.0 && d2 <= 0.0) { ct = 0; q = a; }
    if (d3 >= 0.0 && d4 <= d3) { ct = 0; q = b; }
    if (d6 >= 0.0 && d5 <= d6) { ct = 0; q = c; }
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) { ct = 1; q = a + (d1 / (d1 - d3)) * ab; }
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) { ct = 1; q = a + (d2 / (d2 - d6)) * ac; }
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) { ct = 1; q = b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b); }
    if (va >= 0 && vb >= 0 && vc >= 0) {
        const double denom = 1.0 / (va + vb + vc);
        const V3<double> v = (vb * denom) * ab;
        const V3<double> w = (vc * denom) * ac;
        q = (a + v) + w;
        ct = 2;
    }
    return q;
}

__device__ __forceinline__ bool same_point(V3<double> p, V3<double> c)
{
    return (p.x >= c.x - 1e-9 && p.x <= c.x + 1e-9) && (p.y >= c.y - 1e-9 && p.y <= c.y + 1e-9) &&
           (p.z >= c.z - 1e-9 && p.z <= c.z + 1e-9);
}

// ------------------------------------------------------------------- grid hashing (impl:54-72)
struct Grid {
    double wo_x, wo_y, wo_z, cs;
    uint32_t mask_x, mask_y, mask_z;  // dims - 1 (dims are powers of two)
    uint32_t dim_x, dim_xy;           // dims.x, dims.x * dims.y
    uint32_t base;                    // first bucket of this grid in the bucket table
};

__device__ __forceinline__ void grid_cell(const Grid& g, double x, double y, double z, int& gx, int& gy, int& gz)
{
    gx = (int)floor((x - g.wo_x) / g.cs);
    gy = (int)floor((y - g.wo_y) / g.cs);
    gz = (int)floor((z - g.wo_z) / g.cs);
}

__device__ __forceinline__ uint32_t grid_bucket(const Grid& g, int gx, int gy, int gz)
{
    return g.base + ((uint32_t)gz & g.mask_z) * g.dim_xy + ((uint32_t)gy & g.mask_y) * g.dim_x + ((uint32_t)gx & g.mask_x);
}

// Wave-wide sum of a double (wave64), result valid in every lane.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace xdem

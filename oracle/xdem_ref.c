/*
 * xdem_ref.c -- ORACLE (test infrastructure only; see xdem_ref.h).
 *
 * FP64 restatement of the reference DEM step. Built with -ffp-contract=off so that
 * every multiply and add rounds separately, as the operation order written here does.
 * Line references: impl = include/xdynamics_parallel/xParallelDEM_impl.cuh,
 * def = source/xdynamics_parallel/xParallelDEM_def.cu.
 */
#include "xdem_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { double x, y, z; } v3;
typedef struct { double x, y, z, w; } v4;

static inline v3 mk3(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl3(double s, v3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
static inline v3 div3(v3 a, double s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline v3 neg3(v3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline double dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double len3(v3 a) { return sqrt(dot3(a, a)); }
static inline v3 cross3(v3 a, v3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 ld3(const double* p) { return mk3(p[0], p[1], p[2]); }
static inline v4 ld4(const double* p) { v4 r = {p[0], p[1], p[2], p[3]}; return r; }
static inline void st3(double* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline void st4(double* p, v4 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; p[3] = a.w; }
/* pow(x, 1.5) of the reference, as x*sqrt(x) (see header). */
static inline double pow15(double x) { return x * sqrt(x); }

/* Largest history capacity the restatement accepts (callers clamp to it). */
#define REF_MAX_CAPACITY 64

static int g_threads = 1;
void ref_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int ref_get_threads(void) { return g_threads; }

/* ---------------------------------------------------------------- rotations (impl:8-52) */
static v3 to_global(v3 v, v4 ep)
{
    v3 r0 = mk3(2.0 * (ep.x * ep.x + ep.y * ep.y - 0.5), 2.0 * (ep.y * ep.z - ep.x * ep.w), 2.0 * (ep.y * ep.w + ep.x * ep.z));
    v3 r1 = mk3(2.0 * (ep.y * ep.z + ep.x * ep.w), 2.0 * (ep.x * ep.x + ep.z * ep.z - 0.5), 2.0 * (ep.z * ep.w - ep.x * ep.y));
    v3 r2 = mk3(2.0 * (ep.y * ep.w - ep.x * ep.z), 2.0 * (ep.z * ep.w + ep.x * ep.y), 2.0 * (ep.x * ep.x + ep.w * ep.w - 0.5));
    return mk3(r0.x * v.x + r0.y * v.y + r0.z * v.z,
               r1.x * v.x + r1.y * v.y + r1.z * v.z,
               r2.x * v.x + r2.y * v.y + r2.z * v.z);
}

static v3 to_local(v3 v, v4 ep)
{
    v3 r0 = mk3(2.0 * (ep.x * ep.x + ep.y * ep.y - 0.5), 2.0 * (ep.y * ep.z - ep.x * ep.w), 2.0 * (ep.y * ep.w + ep.x * ep.z));
    v3 r1 = mk3(2.0 * (ep.y * ep.z + ep.x * ep.w), 2.0 * (ep.x * ep.x + ep.z * ep.z - 0.5), 2.0 * (ep.z * ep.w - ep.x * ep.y));
    v3 r2 = mk3(2.0 * (ep.y * ep.w - ep.x * ep.z), 2.0 * (ep.z * ep.w + ep.x * ep.y), 2.0 * (ep.x * ep.x + ep.w * ep.w - 0.5));
    return mk3(r0.x * v.x + r1.x * v.y + r2.x * v.z,
               r0.y * v.x + r1.y * v.y + r2.y * v.z,
               r0.z * v.x + r1.z * v.y + r2.z * v.z);
}

/* omega = 2 L(e) e_dot (impl:34-42) */
static v3 ang_vel(v4 e, v4 d)
{
    v3 o = mk3(-e.y * d.x + e.x * d.y - e.w * d.z + e.z * d.w,
               -e.z * d.x + e.w * d.y + e.x * d.z - e.y * d.w,
               -e.w * d.x - e.z * d.y + e.y * d.z + e.x * d.w);
    return scl3(2.0, o);
}

void ref_to_angular_velocity(const double e[4], const double d[4], double out[3])
{
    st3(out, ang_vel(ld4(e), ld4(d)));
}

/* ------------------------------------------------------------- UCEOM (impl:74-133) */
static v4 uceom(v3 J, v4 ep, v4 ev, v3 n)
{
    /* lhs rows: 2*J_i*G_i(e) for i = 0..2, and e^T */
    double a[4][4];
    a[0][0] = 2.0 * J.x * (-ep.y); a[0][1] = 2.0 * J.x * ep.x;    a[0][2] = 2.0 * J.x * ep.w;    a[0][3] = 2.0 * J.x * (-ep.z);
    a[1][0] = 2.0 * J.y * (-ep.z); a[1][1] = 2.0 * J.y * (-ep.w); a[1][2] = 2.0 * J.y * ep.x;    a[1][3] = 2.0 * J.y * ep.y;
    a[2][0] = 2.0 * J.z * (-ep.w); a[2][1] = 2.0 * J.z * ep.z;    a[2][2] = 2.0 * J.z * (-ep.y); a[2][3] = 2.0 * J.z * ep.x;
    a[3][0] = ep.x;                a[3][1] = ep.y;                a[3][2] = ep.z;                a[3][3] = ep.w;
    /* T = J G(e) e_dot */
    v3 T = mk3(J.x * (-ep.y * ev.x + ep.x * ev.y + ep.w * ev.z - ep.z * ev.w),
               J.y * (-ep.z * ev.x - ep.w * ev.y + ep.x * ev.z + ep.y * ev.w),
               J.z * (-ep.w * ev.x + ep.z * ev.y - ep.y * ev.z + ep.x * ev.w));
    /* 4 G(e_dot)^T T */
    v4 L0;
    L0.x = 4.0 * (-ev.y * T.x - ev.z * T.y - ev.w * T.z);
    L0.y = 4.0 * (ev.x * T.x - ev.w * T.y + ev.z * T.z);
    L0.z = 4.0 * (ev.w * T.x + ev.x * T.y - ev.y * T.z);
    L0.w = 4.0 * (-ev.z * T.x + ev.y * T.y + ev.x * T.z);
    /* G(e) L0 */
    v3 LH = mk3(-ep.y * L0.x + ep.x * L0.y + ep.w * L0.z - ep.z * L0.w,
                -ep.z * L0.x - ep.w * L0.y + ep.x * L0.z + ep.y * L0.w,
                -ep.w * L0.x + ep.z * L0.y - ep.y * L0.z + ep.x * L0.w);
    double evev = ev.x * ev.x + ev.y * ev.y + ev.z * ev.z + ev.w * ev.w;
    double rhs[4] = {n.x - LH.x, n.y - LH.y, n.z - LH.z, 0.0 - evev};

    /* Determinant and adjugate by cofactors, terms in the reference's order (impl:98-126). */
#define A(i, j) a[i][j]
    double det =
        A(0,0)*A(1,1)*A(2,2)*A(3,3) + A(0,0)*A(1,2)*A(2,3)*A(3,1) + A(0,0)*A(1,3)*A(2,1)*A(3,2) -
        A(0,0)*A(1,3)*A(2,2)*A(3,1) - A(0,0)*A(1,2)*A(2,1)*A(3,3) - A(0,0)*A(1,1)*A(2,3)*A(3,2) -
        A(0,1)*A(1,0)*A(2,2)*A(3,3) - A(0,2)*A(1,0)*A(2,3)*A(3,1) - A(0,3)*A(1,0)*A(2,1)*A(3,2) +
        A(0,3)*A(1,0)*A(2,2)*A(3,1) + A(0,2)*A(1,0)*A(2,1)*A(3,3) + A(0,1)*A(1,0)*A(2,3)*A(3,2) +
        A(0,1)*A(1,2)*A(2,0)*A(3,3) + A(0,2)*A(1,3)*A(2,0)*A(3,1) + A(0,3)*A(1,1)*A(2,0)*A(3,2) -
        A(0,3)*A(1,2)*A(2,0)*A(3,1) - A(0,2)*A(1,1)*A(2,0)*A(3,3) - A(0,1)*A(1,3)*A(2,0)*A(3,2) -
        A(0,1)*A(1,2)*A(2,3)*A(3,0) - A(0,2)*A(1,3)*A(2,1)*A(3,0) - A(0,3)*A(1,1)*A(2,2)*A(3,0) +
        A(0,3)*A(1,2)*A(2,1)*A(3,0) + A(0,2)*A(1,1)*A(2,3)*A(3,0) + A(0,1)*A(1,3)*A(2,2)*A(3,0);
    double o[4][4];
    o[0][0] =  A(1,1)*A(2,2)*A(3,3) + A(1,2)*A(2,3)*A(3,1) + A(1,3)*A(2,1)*A(3,2) - A(1,3)*A(2,2)*A(3,1) - A(1,2)*A(2,1)*A(3,3) - A(1,1)*A(2,3)*A(3,2);
    o[0][1] = -A(0,1)*A(2,2)*A(3,3) - A(0,2)*A(2,3)*A(3,1) - A(0,3)*A(2,1)*A(3,2) + A(0,3)*A(2,2)*A(3,1) + A(0,2)*A(2,1)*A(3,3) + A(0,1)*A(2,3)*A(3,2);
    o[0][2] =  A(0,1)*A(1,2)*A(3,3) + A(0,2)*A(1,3)*A(3,1) + A(0,3)*A(1,1)*A(3,2) - A(0,3)*A(1,2)*A(3,1) - A(0,2)*A(1,1)*A(3,3) - A(0,1)*A(1,3)*A(3,2);
    o[0][3] = -A(0,1)*A(1,2)*A(2,3) - A(0,2)*A(1,3)*A(2,1) - A(0,3)*A(1,1)*A(2,2) + A(0,3)*A(1,2)*A(2,1) + A(0,2)*A(1,1)*A(2,3) + A(0,1)*A(1,3)*A(2,2);
    o[1][0] = -A(1,0)*A(2,2)*A(3,3) - A(1,2)*A(2,3)*A(3,0) - A(1,3)*A(2,0)*A(3,2) + A(1,3)*A(2,2)*A(3,0) + A(1,2)*A(2,0)*A(3,3) + A(1,0)*A(2,3)*A(3,2);
    o[1][1] =  A(0,0)*A(2,2)*A(3,3) + A(0,2)*A(2,3)*A(3,0) + A(0,3)*A(2,0)*A(3,2) - A(0,3)*A(2,2)*A(3,0) - A(0,2)*A(2,0)*A(3,3) - A(0,0)*A(2,3)*A(3,2);
    o[1][2] = -A(0,0)*A(1,2)*A(3,3) - A(0,2)*A(1,3)*A(3,0) - A(0,3)*A(1,0)*A(3,2) + A(0,3)*A(1,2)*A(3,0) + A(0,2)*A(1,0)*A(3,3) + A(0,0)*A(1,3)*A(3,2);
    o[1][3] =  A(0,0)*A(1,2)*A(2,3) + A(0,2)*A(1,3)*A(2,0) + A(0,3)*A(1,0)*A(2,2) - A(0,3)*A(1,2)*A(2,0) - A(0,2)*A(1,0)*A(2,3) - A(0,0)*A(1,3)*A(2,2);
    o[2][0] =  A(1,0)*A(2,1)*A(3,3) + A(1,1)*A(2,3)*A(3,0) + A(1,3)*A(2,0)*A(3,1) - A(1,3)*A(2,1)*A(3,0) - A(1,1)*A(2,0)*A(3,3) - A(1,0)*A(2,3)*A(3,1);
    o[2][1] = -A(0,0)*A(2,1)*A(3,3) - A(0,1)*A(2,3)*A(3,0) - A(0,3)*A(2,0)*A(3,1) + A(0,3)*A(2,1)*A(3,0) + A(0,1)*A(2,0)*A(3,3) + A(0,0)*A(2,3)*A(3,1);
    o[2][2] =  A(0,0)*A(1,1)*A(3,3) + A(0,1)*A(1,3)*A(3,0) + A(0,3)*A(1,0)*A(3,1) - A(0,3)*A(1,1)*A(3,0) - A(0,1)*A(1,0)*A(3,3) - A(0,0)*A(1,3)*A(3,1);
    o[2][3] = -A(0,0)*A(1,1)*A(2,3) - A(0,1)*A(1,3)*A(2,0) - A(0,3)*A(1,0)*A(2,1) + A(0,3)*A(1,1)*A(2,0) + A(0,1)*A(1,0)*A(2,3) + A(0,0)*A(1,3)*A(2,1);
    o[3][0] = -A(1,0)*A(2,1)*A(3,2) - A(1,1)*A(2,2)*A(3,0) - A(1,2)*A(2,0)*A(3,1) + A(1,2)*A(2,1)*A(3,0) + A(1,1)*A(2,0)*A(3,2) + A(1,0)*A(2,2)*A(3,1);
    o[3][1] =  A(0,0)*A(2,1)*A(3,2) + A(0,1)*A(2,2)*A(3,0) + A(0,2)*A(2,0)*A(3,1) - A(0,2)*A(2,1)*A(3,0) - A(0,1)*A(2,0)*A(3,2) - A(0,0)*A(2,2)*A(3,1);
    o[3][2] = -A(0,0)*A(1,1)*A(3,2) - A(0,1)*A(1,2)*A(3,0) - A(0,2)*A(1,0)*A(3,1) + A(0,2)*A(1,1)*A(3,0) + A(0,1)*A(1,0)*A(3,2) + A(0,0)*A(1,2)*A(3,1);
    o[3][3] =  A(0,0)*A(1,1)*A(2,2) + A(0,1)*A(1,2)*A(2,0) + A(0,2)*A(1,0)*A(2,1) - A(0,2)*A(1,1)*A(2,0) - A(0,1)*A(1,0)*A(2,2) - A(0,0)*A(1,2)*A(2,1);
#undef A
    double m = 1.0 / det;
    v4 r;
    r.x = m * (o[0][0] * rhs[0] + o[0][1] * rhs[1] + o[0][2] * rhs[2] + o[0][3] * rhs[3]);
    r.y = m * (o[1][0] * rhs[0] + o[1][1] * rhs[1] + o[1][2] * rhs[2] + o[1][3] * rhs[3]);
    r.z = m * (o[2][0] * rhs[0] + o[2][1] * rhs[1] + o[2][2] * rhs[2] + o[2][3] * rhs[3]);
    r.w = m * (o[3][0] * rhs[0] + o[3][1] * rhs[1] + o[3][2] * rhs[2] + o[3][3] * rhs[3]);
    return r;
}

void ref_calculate_uceom(const double J[3], const double ep[4], const double ev[4],
                         const double n_prime[3], double out[4])
{
    st4(out, uceom(ld3(J), ld4(ep), ld4(ev), ld3(n_prime)));
}

/* ------------------------------------------------------- Velocity Verlet (impl:135-222) */
void ref_vv_update_position(const ref_dem_parameters* p, double* pos, double* ep,
                            const double* vel, const double* ev, const double* acc,
                            const double* ea, unsigned int np)
{
    const double dt = p->dt, h = p->half2dt;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long i = 0; i < (long)np; i++) {
        v4 e = ld4(ep + 4 * i), d = ld4(ev + 4 * i), a = ld4(ea + 4 * i);
        v4 ne;
        ne.x = e.x + dt * d.x + h * a.x;
        ne.y = e.y + dt * d.y + h * a.y;
        ne.z = e.z + dt * d.z + h * a.z;
        ne.w = e.w + dt * d.w + h * a.w;
        double l = sqrt(ne.x * ne.x + ne.y * ne.y + ne.z * ne.z + ne.w * ne.w);
        ne.x = ne.x / l; ne.y = ne.y / l; ne.z = ne.z / l; ne.w = ne.w / l;
        st4(ep + 4 * i, ne);
        v3 v = ld3(vel + 3 * i), ac = ld3(acc + 3 * i);
        v3 dp = mk3(dt * v.x + h * ac.x, dt * v.y + h * ac.y, dt * v.z + h * ac.z);
        pos[4 * i + 0] += dp.x;
        pos[4 * i + 1] += dp.y;
        pos[4 * i + 2] += dp.z;
    }
}

void ref_vv_update_velocity(const ref_dem_parameters* p, const ref_quirks* q,
                            double* vel, double* acc, const double* ep, double* ev,
                            double* ea, double* force, double* moment,
                            const double* mass, const double* iner, unsigned int np)
{
    const double dt = p->dt;
    const v3 g = ld3(p->gravity);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long i = 0; i < (long)np; i++) {
        double m = mass[i];
        v3 F = ld3(force + 3 * i);
        v4 e = ld4(ep + 4 * i), av = ld4(ev + 4 * i), aa = ld4(ea + 4 * i);
        v3 J = ld3(iner + 3 * i);
        v3 np_ = to_local(ld3(moment + 3 * i), e);
        v4 mea = uceom(J, e, av, np_);
        double inv_m = 1.0 / m;
        v3 a = mk3(inv_m * (F.x + m * g.x), inv_m * (F.y + m * g.y), inv_m * (F.z + m * g.z));
        double hdt = 0.5 * dt;
        if (q->velocity_not_updated) {
            /* impl:211-221: velocity stored unchanged; e_dot += dt/2 (ea_old + ea_new) */
            av.x = av.x + hdt * (aa.x + mea.x);
            av.y = av.y + hdt * (aa.y + mea.y);
            av.z = av.z + hdt * (aa.z + mea.z);
            av.w = av.w + hdt * (aa.w + mea.w);
        } else {
            /* xIntegratorVV.cpp:143-149 */
            v3 v = ld3(vel + 3 * i), ao = ld3(acc + 3 * i);
            v = mk3(v.x + hdt * ao.x, v.y + hdt * ao.y, v.z + hdt * ao.z);
            av.x = av.x + hdt * aa.x; av.y = av.y + hdt * aa.y;
            av.z = av.z + hdt * aa.z; av.w = av.w + hdt * aa.w;
            v = mk3(v.x + hdt * a.x, v.y + hdt * a.y, v.z + hdt * a.z);
            av.x = av.x + hdt * mea.x; av.y = av.y + hdt * mea.y;
            av.z = av.z + hdt * mea.z; av.w = av.w + hdt * mea.w;
            st3(vel + 3 * i, v);
        }
        st4(ev + 4 * i, av);
        st3(acc + 3 * i, a);
        st4(ea + 4 * i, mea);
        st3(force + 3 * i, mk3(0.0, 0.0, 0.0));
        st3(moment + 3 * i, mk3(0.0, 0.0, 0.0));
    }
}

/* ------------------------------------------------------------ hash (impl:54-72, 290-302) */
static inline unsigned int grid_hash(const ref_dem_parameters* p, int gx, int gy, int gz)
{
    unsigned int x = (unsigned int)gx & (p->grid_size[0] - 1);
    unsigned int y = (unsigned int)gy & (p->grid_size[1] - 1);
    unsigned int z = (unsigned int)gz & (p->grid_size[2] - 1);
    return (z * p->grid_size[1]) * p->grid_size[0] + y * p->grid_size[0] + x;
}

static inline void grid_pos(const ref_dem_parameters* p, v3 x, int* g)
{
    g[0] = (int)floor((x.x - p->world_origin[0]) / p->cell_size);
    g[1] = (int)floor((x.y - p->world_origin[1]) / p->cell_size);
    g[2] = (int)floor((x.z - p->world_origin[2]) / p->cell_size);
}

void ref_calculate_hash_and_index(const ref_dem_parameters* p, unsigned int* hash,
                                  unsigned int* index, const double* pos,
                                  unsigned int sid, unsigned int np)
{
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long i = 0; i < (long)np; i++) {
        int g[3];
        grid_pos(p, ld3(pos + 4 * i), g);
        hash[sid + i] = grid_hash(p, g[0], g[1], g[2]);
        index[sid + i] = sid + (unsigned int)i;
    }
}

/* --------------------------------------------------- sort + cell bounds (def:94-124, impl:317-360) */
static int cmp_u64(const void* a, const void* b)
{
    unsigned long long x = *(const unsigned long long*)a, y = *(const unsigned long long*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

void ref_reorder_data_and_find_cell_start(unsigned int* hash, unsigned int* index,
                                          unsigned int* cstart, unsigned int* cend,
                                          unsigned int* sorted_index, unsigned int n,
                                          unsigned int ncell)
{
    /* A stable sort by hash keeps equal hashes in input order; the input index
     * array ascends (impl:301), so ordering by (hash, index) is the same result. */
    unsigned long long* kv = (unsigned long long*)malloc(sizeof(unsigned long long) * (n ? n : 1));
    for (unsigned int i = 0; i < n; i++)
        kv[i] = ((unsigned long long)hash[i] << 32) | (unsigned long long)i;
    qsort(kv, n, sizeof(unsigned long long), cmp_u64);
    unsigned int* idx_copy = (unsigned int*)malloc(sizeof(unsigned int) * (n ? n : 1));
    memcpy(idx_copy, index, sizeof(unsigned int) * n);
    for (unsigned int i = 0; i < n; i++) {
        hash[i] = (unsigned int)(kv[i] >> 32);
        index[i] = idx_copy[(unsigned int)(kv[i] & 0xffffffffu)];
    }
    free(idx_copy);
    free(kv);
    memset(cstart, 0xff, sizeof(unsigned int) * ncell);
    memset(cend, 0, sizeof(unsigned int) * ncell);
    for (unsigned int i = 0; i < n; i++) {
        unsigned int h = hash[i];
        if (i == 0 || h != hash[i - 1]) {
            cstart[h] = i;
            if (i > 0) cend[hash[i - 1]] = i;
        }
        if (i == n - 1) cend[h] = i + 1;
        sorted_index[i] = index[i];
    }
}

/* ------------------------------------------------------------- contact constants (impl:362-386) */
typedef struct { double kn, vn, ks, vs, mu, mu_s, ms, eq_r, eq_m; } fconst;

static fconst get_constant(int model, double ir, double jr, double im, double jm,
                           const ref_contact_property* cp)
{
    fconst c;
    double iE = cp->Ei, jE = cp->Ej, ip = cp->pri, jp = cp->prj;
    double Meq = jm ? (im * jm) / (im + jm) : im;
    double Req = jr ? (ir * jr) / (ir + jr) : ir;
    double Eeq = (iE * jE) / (iE * (1 - jp * jp) + jE * (1 - ip * ip));
    double Seq = (2.0 * (2.0 - ip) * (1.0 + ip) / iE) + (2.0 * (2.0 - jp) * (1.0 + jp) / jE);
    Seq = 1.0 / Seq;
    double lne = log(cp->rest);
    double beta = -lne * sqrt(1.0 / (lne * lne + M_PI * M_PI));
    c.eq_m = Meq;
    c.eq_r = Req;
    c.kn = (4.0 / 3.0) * Eeq * sqrt(Req);
    c.vn = beta;
    c.ks = model == 1 ? 8.0 * Seq : c.kn * cp->sratio;
    c.vs = model == 1 ? 2.0 * sqrt(5.0 / 6.0) * beta : c.vn * cp->sratio;
    c.mu = cp->fric;
    c.mu_s = cp->s_fric;
    c.ms = cp->rfric;
    return c;
}

/* ------------------------------------------------------------ cohesion (impl:388-430) */
double ref_cohesion_force(double ri, double rj, const ref_contact_property* cp, double Fn)
{
    double cf = 0.0;
    double coh = cp->coh;
    if (coh) {
        double Req = rj ? (ri * rj) / (ri + rj) : ri;
        double Eeq = 1.0 / (((1.0 - cp->pri * cp->pri) / cp->Ei) + ((1.0 - cp->prj * cp->prj) / cp->Ej));
        double c0 = 3.0 * coh * M_PI * Req;
        double eq = 2.0 * c0 * Fn + c0 * c0;
        if (eq <= 0) Fn = -0.5 * c0;
        double a3 = (3.0 * Req) * (Fn + c0 + sqrt(fabs(2.0 * c0 * Fn + c0 * c0))) / (4.0 * Eeq);
        cf = -sqrt(8.0 * M_PI * coh * Eeq * a3);
    }
    return cf;
}

double ref_cohesion_limit_depth(double ir, double jr, const ref_contact_property* cp)
{
    double Req = jr ? (ir * jr) / (ir + jr) : ir;
    double Eeq = 1.0 / (((1.0 - cp->pri * cp->pri) / cp->Ei) + ((1.0 - cp->prj * cp->prj) / cp->Ej));
    double c1 = (M_PI * M_PI * cp->coh * cp->coh * Req) / (Eeq * Eeq);
    return -(3.0 / 4.0) * pow(c1, 1.0 / 3.0);
}

static double jkr_separation_force(double ir, double jr, double coh)
{
    double Req = jr ? (ir * jr) / (ir + jr) : ir;
    return -(3.0 / 2.0) * M_PI * coh * Req;
}

/* ------------------------------------------------- force law (impl:432-474; xContact.cpp:210-264) */
static void force_law(int model, int gpu_law, double dt, fconst c, double ir, double jr,
                      const ref_contact_property* cp, double cdist, double* ds_io, double* dots_io,
                      v3 dv, v3 unit, v3* Ft, v3* Fn)
{
    double fsn = -c.kn * pow15(cdist);
    double dvu = dot3(dv, unit);
    double fdn;
    if (model == 1 && !gpu_law)
        fdn = c.vn * (2.0 * sqrt(c.eq_m * c.kn)) * dvu;   /* xContact.cpp:244 */
    else
        fdn = c.vn * dvu;                                  /* impl:438, 460; xContact.cpp:216 */
    double fca = -ref_cohesion_force(ir, jr, cp, fsn + fdn);
    /* summation order differs between the two texts: CPU (fsn + fsd + fco),
     * kernel (fsn + fca + fdn), for both models */
    double fnm = !gpu_law ? (fsn + fdn + fca) : (fsn + fca + fdn);
    *Fn = scl3(fnm, unit);
    v3 e = sub3(dv, scl3(dvu, unit));
    double mag_e = len3(e);
    if (mag_e) {
        v3 sh = div3(e, mag_e);
        double s_dot = dot3(dv, sh);
        double ds = *ds_io + dt * (s_dot + *dots_io);
        *ds_io = ds;
        *dots_io = s_dot;
        double ft1;
        if (model == 1) {
            double St = c.ks * sqrt(c.eq_r * cdist);
            if (gpu_law)
                ft1 = St * ds + c.vs * sqrt(St * c.eq_m);              /* impl:450 */
            else
                ft1 = St * ds + c.vs * sqrt(St * c.eq_m) * s_dot;      /* xContact.cpp:257 */
        } else {
            ft1 = c.ks * ds + c.vs * s_dot;                            /* impl:472; xContact.cpp:229 */
        }
        double ft2 = c.mu * len3(*Fn);
        *Ft = scl3(fmin(ft1, ft2), sh);
    }
}

void ref_pair_force(const ref_dem_parameters* p, const ref_quirks* q,
                    const ref_contact_property* cp, double ir, double jr, double im, double jm,
                    double cdist, const double dv[3], const double unit[3],
                    double io_sd[2], double io_ft[3], double fn_out[3])
{
    fconst c = get_constant(p->contact_model, ir, jr, im, jm, cp);
    v3 Ft = ld3(io_ft), Fn;
    force_law(p->contact_model, q->gpu_force_law, p->dt, c, ir, jr, cp, cdist, &io_sd[0], &io_sd[1],
              ld3(dv), ld3(unit), &Ft, &Fn);
    st3(io_ft, Ft);
    st3(fn_out, Fn);
}

/* Optional per-sphere output of the sum of sphere-contact force magnitudes (the
 * normaliser of the mixed-precision error metric); NULL = off. */
static double* g_pair_mag = NULL;
void ref_set_pair_magnitude_output(double* out) { g_pair_mag = out; }

/* ------------------------------------------------------- sphere-sphere (impl:613-733) */
unsigned int ref_calculate_p2p(const ref_dem_parameters* p, const ref_quirks* q,
                               const double* pos, const double* ep, const double* vel,
                               const double* ev, double* force, double* moment,
                               const double* mass, double* tmax, double* rres,
                               unsigned int* pair_count, unsigned int* pair_id, double* tsd,
                               const unsigned int* sorted_index, const unsigned int* cstart,
                               const unsigned int* cend, const ref_contact_property* cp,
                               unsigned int np, unsigned int capacity)
{
    unsigned int overflow = 0;
    const int model = p->contact_model;
    if (capacity > REF_MAX_CAPACITY) capacity = REF_MAX_CAPACITY;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads) reduction(+:overflow)
    for (long ii = 0; ii < (long)np; ii++) {
        unsigned int id = (unsigned int)ii;
        unsigned int sid = id * capacity;
        unsigned int old_count = pair_count[id];
        if (old_count > capacity) old_count = capacity;
        /* previous slots, read before this particle writes any (impl:625-633) */
        unsigned int p_id[REF_MAX_CAPACITY];
        double p_sd[2 * REF_MAX_CAPACITY];
        for (unsigned int s = 0; s < capacity; s++) {
            p_id[s] = pair_id[sid + s];
            p_sd[2 * s] = tsd[2 * (sid + s)];
            p_sd[2 * s + 1] = tsd[2 * (sid + s) + 1];
        }
        v4 ipos = ld4(pos + 4 * id);
        v3 ip3 = mk3(ipos.x, ipos.y, ipos.z);
        v3 ivel = ld3(vel + 3 * id);
        v3 iom = ang_vel(ld4(ep + 4 * id), ld4(ev + 4 * id));
        int g[3];
        grid_pos(p, ip3, g);
        double ir = ipos.w, im = mass[id];
        v3 Ft = mk3(0, 0, 0), Fn = mk3(0, 0, 0);
        v3 sumF = mk3(0, 0, 0), sumM = mk3(0, 0, 0);
        unsigned int n_new = 0, n_seen = 0;
        double res = 0.0, mag = 0.0;
        v3 tma = mk3(0, 0, 0);
        for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
        for (int x = -1; x <= 1; x++) {
            unsigned int h = grid_hash(p, g[0] + x, g[1] + y, g[2] + z);
            unsigned int start = cstart[h];
            if (start == 0xffffffffu) continue;
            unsigned int end = cend[h];
            for (unsigned int j = start; j < end; j++) {
                unsigned int k = sorted_index[j];
                if (id == k || k >= np) continue;
                v4 jpos = ld4(pos + 4 * k);
                v3 jvel = ld3(vel + 3 * k);
                v3 jom = ang_vel(ld4(ep + 4 * k), ld4(ev + 4 * k));
                double jr = jpos.w, jm = mass[k];
                v3 rp = mk3(jpos.x - ipos.x, jpos.y - ipos.y, jpos.z - ipos.z);
                double dist = len3(rp);
                double cdist = (ir + jr) - dist;
                double coh_s = 0;
                if (cp->coh) coh_s = ref_cohesion_limit_depth(ir, jr, cp);
                double sd0 = 0.0, sd1 = 0.0;
                v3 unit = div3(rp, dist);
                v3 cpt = add3(ip3, scl3(ir, unit));
                if (cdist > 0) {
                    for (unsigned int s = 0; s < old_count; s++)
                        if (p_id[s] == k) { sd0 = p_sd[2 * s]; sd1 = p_sd[2 * s + 1]; break; }
                    v3 dcpr = sub3(cpt, ip3);
                    v3 dcpr_j = sub3(cpt, mk3(jpos.x, jpos.y, jpos.z));
                    v3 rv = sub3(add3(jvel, cross3(jom, dcpr_j)), add3(ivel, cross3(iom, dcpr)));
                    fconst c = get_constant(model, ir, jr, im, jm, cp);
                    if (!q->stale_ft) Ft = mk3(0, 0, 0);
                    force_law(model, q->gpu_force_law, p->dt, c, ir, jr, cp, cdist, &sd0, &sd1, rv, unit, &Ft, &Fn);
                    /* rolling accumulation (impl:476-483) */
                    v3 fnt = add3(Fn, Ft);
                    tma = add3(tma, cross3(dcpr, fnt));
                    double Rij = jr ? ir * jr / (ir + jr) : ir;
                    res += Rij * cp->rfric * len3(Fn);
                    sumF = add3(sumF, fnt);
                    sumM = add3(sumM, cross3(dcpr, fnt));
                    mag += len3(fnt);
                    n_seen++;
                    if (n_new < capacity) {
                        tsd[2 * (sid + n_new)] = sd0;
                        tsd[2 * (sid + n_new) + 1] = sd1;
                        pair_id[sid + n_new] = k;
                        n_new++;
                    }
                } else if (cdist <= 0 && fabs(cdist) < fabs(coh_s)) {
                    double f = jkr_separation_force(ir, jr, cp->coh);
                    double cf = ref_cohesion_force(ir, jr, cp, f);
                    sumF = sub3(sumF, scl3(cf, unit));
                    mag += fabs(cf);
                    n_seen++;
                    if (n_new < capacity) {
                        tsd[2 * (sid + n_new)] = sd0;
                        tsd[2 * (sid + n_new) + 1] = sd1;
                        pair_id[sid + n_new] = k;
                        n_new++;
                    }
                }
            }
        }
        double* F = force + 3 * id;
        double* M = moment + 3 * id;
        F[0] += sumF.x; F[1] += sumF.y; F[2] += sumF.z;
        M[0] += sumM.x; M[1] += sumM.y; M[2] += sumM.z;
        if (n_seen > capacity) overflow++;
        pair_count[id] = n_new;
        tmax[3 * id + 0] += tma.x; tmax[3 * id + 1] += tma.y; tmax[3 * id + 2] += tma.z;
        rres[id] += res;
        if (g_pair_mag) g_pair_mag[id] = mag;
    }
    return overflow;
}

/* ---------------------------------------------------------- wall detection (impl:735-815) */
static double plane_detection(const ref_plane_info* pe, v3 xp, v3 wp, v3* u, double r)
{
    double a_l1 = (wp.x - pe->l1) * (wp.x - pe->l1);
    double b_l2 = (wp.y - pe->l2) * (wp.y - pe->l2);
    double sqa = wp.x * wp.x, sqb = wp.y * wp.y, sqc = wp.z * wp.z, sqr = r * r;
    v3 xw = ld3(pe->xw), w2 = ld3(pe->w2), w3 = ld3(pe->w3), w4 = ld3(pe->w4);
    /* corner regions: unit points from the corner to the sphere (as written) */
    if (wp.x < 0 && wp.y < 0 && (sqa + sqb + sqc) < sqr) {
        v3 X = sub3(xp, xw); double h = len3(X); *u = div3(X, h); return r - h;
    } else if (wp.x > pe->l1 && wp.y < 0 && (a_l1 + sqb + sqc) < sqr) {
        v3 X = sub3(xp, w2); double h = len3(X); *u = div3(X, h); return r - h;
    } else if (wp.x > pe->l1 && wp.y > pe->l2 && (a_l1 + b_l2 + sqc) < sqr) {
        v3 X = sub3(xp, w3); double h = len3(X); *u = div3(X, h); return r - h;
    } else if (wp.x < 0 && wp.y > pe->l2 && (sqa + b_l2 + sqc) < sqr) {
        v3 X = sub3(xp, w4); double h = len3(X); *u = div3(X, h); return r - h;
    }
    /* edge regions, then the face */
    if ((wp.x > 0 && wp.x < pe->l1) && wp.y < 0 && (sqb + sqc) < sqr) {
        v3 X = sub3(xp, xw), e = sub3(w2, xw);
        v3 us = div3(e, len3(e));
        v3 hs = sub3(X, scl3(dot3(X, us), us));
        double h = len3(hs); *u = div3(neg3(hs), h); return r - h;
    } else if ((wp.x > 0 && wp.x < pe->l1) && wp.y > pe->l2 && (b_l2 + sqc) < sqr) {
        v3 X = sub3(xp, w4), e = sub3(w3, w4);
        v3 us = div3(e, len3(e));
        v3 hs = sub3(X, scl3(dot3(X, us), us));
        double h = len3(hs); *u = div3(neg3(hs), h); return r - h;
    } else if ((wp.x > 0 && wp.y < pe->l2) && wp.x < 0 && (sqr + sqc) < sqr) {
        /* unreachable as written (wp.x > 0 && wp.x < 0), impl:787-795 */
        v3 X = sub3(xp, xw), e = sub3(w4, xw);
        v3 us = div3(e, len3(e));
        v3 hs = sub3(X, scl3(dot3(X, us), us));
        double h = len3(hs); *u = div3(neg3(hs), h); return r - h;
    } else if ((wp.x > 0 && wp.y < pe->l2) && wp.x > pe->l1 && (a_l1 + sqc) < sqr) {
        v3 X = sub3(xp, w2), e = sub3(w3, w2);
        v3 us = div3(e, len3(e));
        v3 hs = sub3(X, scl3(dot3(X, us), us));
        double h = len3(hs); *u = div3(neg3(hs), h); return r - h;
    } else if ((wp.x > 0 && wp.x < pe->l1) && (wp.y > 0 && wp.y < pe->l2)) {
        v3 dp = sub3(xp, xw);
        v3 uw = ld3(pe->uw);
        v3 uu = div3(uw, len3(uw));
        *u = neg3(uu);
        return r - fabs(dot3(dp, *u));
    }
    return 0;
}

double ref_plane_detection(const ref_plane_info* pe, const double xp[3], double r, double u_out[3])
{
    v3 x = ld3(xp);
    v3 dp = sub3(x, ld3(pe->xw));
    v3 wp = mk3(dot3(dp, ld3(pe->u1)), dot3(dp, ld3(pe->u2)), dot3(dp, ld3(pe->uw)));
    v3 u = mk3(0, 0, 0);
    double c = plane_detection(pe, x, wp, &u, r);
    st3(u_out, u);
    return c;
}

/* ------------------------------------------------------------ wall contact (impl:950-1080) */
unsigned int ref_plane_contact_force(const ref_dem_parameters* p, const ref_quirks* q,
                                     const ref_plane_info* plane, const double* dbi,
                                     const ref_contact_property* cp,
                                     double* fx, double* fy, double* fz,
                                     double* mx, double* my, double* mz,
                                     const double* pos, const double* ep, const double* vel,
                                     const double* ev, double* force, double* moment,
                                     const double* mass, double* tmax, double* rres,
                                     unsigned int* pair_count, unsigned int* pair_id,
                                     double* tsd, unsigned int np, unsigned int capacity)
{
    unsigned int overflow = 0;
    const int model = p->contact_model;
    const v3 bpos = mk3(dbi[REF_PX], dbi[REF_PY], dbi[REF_PZ]);
    const v3 bvel = mk3(dbi[REF_VX], dbi[REF_VY], dbi[REF_VZ]);
    v4 be = {dbi[REF_E0], dbi[REF_E1], dbi[REF_E2], dbi[REF_E3]};
    v4 bd = {dbi[REF_ED0], dbi[REF_ED1], dbi[REF_ED2], dbi[REF_ED3]};
    const v3 oj = ang_vel(be, bd);
    if (capacity > REF_MAX_CAPACITY) capacity = REF_MAX_CAPACITY;
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+:overflow)
    for (long ii = 0; ii < (long)np; ii++) {
        unsigned int id = (unsigned int)ii;
        unsigned int sid = id * capacity;
        unsigned int p_id[REF_MAX_CAPACITY]; double p_sd[2 * REF_MAX_CAPACITY];
        for (unsigned int s = 0; s < capacity; s++) {
            p_id[s] = pair_id[sid + s];
            p_sd[2 * s] = tsd[2 * (sid + s)];
            p_sd[2 * s + 1] = tsd[2 * (sid + s) + 1];
        }
        double bf[6] = {0, 0, 0, 0, 0, 0};
        unsigned int old_count = pair_count[id];
        double m = mass[id];
        v4 ipos = ld4(pos + 4 * id);
        v3 ip3 = mk3(ipos.x, ipos.y, ipos.z);
        double r = ipos.w;
        unsigned int k = plane->id;
        v3 ivel = ld3(vel + 3 * id);
        v3 iom = ang_vel(ld4(ep + 4 * id), ld4(ev + 4 * id));
        v3 Fn = mk3(0, 0, 0), Ft = mk3(0, 0, 0);
        v3 sumF = mk3(0, 0, 0), sumM = mk3(0, 0, 0);
        unsigned int n_new = k ? old_count : 0;
        if (n_new > capacity) n_new = capacity;
        double res = 0.0;
        v3 tma = mk3(0, 0, 0);
        double coh_s = 0;
        v3 dp = sub3(ip3, ld3(plane->xw));
        v3 unit = mk3(0, 0, 0);
        v3 wp = mk3(dot3(dp, ld3(plane->u1)), dot3(dp, ld3(plane->u2)), dot3(dp, ld3(plane->uw)));
        double cdist = plane_detection(plane, ip3, wp, &unit, r);
        if (cp->coh) coh_s = ref_cohesion_limit_depth(r, 0, cp);
        double sd0 = 0.0, sd1 = 0.0;
        v3 cpt = add3(ip3, scl3(r, unit));
        v3 dcpr_j = sub3(cpt, bpos);
        int stored = 0;
        if (cdist > 0) {
            for (unsigned int s = 0; s < capacity; s++)
                if (p_id[s] == k) { sd0 = p_sd[2 * s]; sd1 = p_sd[2 * s + 1]; break; }
            v3 dcpr = sub3(cpt, ip3);
            v3 dv = sub3(add3(bvel, cross3(oj, dcpr_j)), add3(ivel, cross3(iom, dcpr)));
            fconst c = get_constant(model, r, 0.0, m, 0.0, cp);
            force_law(model, q->gpu_force_law, p->dt, c, r, 0, cp, cdist, &sd0, &sd1, dv, unit, &Ft, &Fn);
            v3 mf = add3(Fn, Ft);
            tma = add3(tma, cross3(dcpr, mf));
            res += r * cp->rfric * len3(Fn);
            sumF = add3(sumF, mf);
            sumM = add3(sumM, cross3(dcpr, mf));
            bf[0] += -mf.x; bf[1] += -mf.y; bf[2] += -mf.z;
            v3 bm = neg3(cross3(dcpr_j, mf));
            bf[3] += bm.x; bf[4] += bm.y; bf[5] += bm.z;
            stored = 1;
        } else if (cdist <= 0 && fabs(cdist) < fabs(coh_s)) {
            double f = jkr_separation_force(r, 0, cp->coh);
            double cf = ref_cohesion_force(r, 0, cp, f);
            v3 mf = scl3(-cf, unit);
            sumF = add3(sumF, mf);
            bf[0] += -mf.x; bf[1] += -mf.y; bf[2] += -mf.z;
            v3 bm = neg3(cross3(dcpr_j, mf));
            bf[3] += bm.x; bf[4] += bm.y; bf[5] += bm.z;
            stored = 1;
        }
        if (stored) {
            if (n_new < capacity) {
                tsd[2 * (sid + n_new)] = sd0;
                tsd[2 * (sid + n_new) + 1] = sd1;
                pair_id[sid + n_new] = k;
                n_new++;
            } else {
                overflow++;
            }
        }
        if (fx) { fx[id] = bf[0]; fy[id] = bf[1]; fz[id] = bf[2]; mx[id] = bf[3]; my[id] = bf[4]; mz[id] = bf[5]; }
        double* F = force + 3 * id;
        double* M = moment + 3 * id;
        F[0] += sumF.x; F[1] += sumF.y; F[2] += sumF.z;
        M[0] += sumM.x; M[1] += sumM.y; M[2] += sumM.z;
        pair_count[id] = n_new;
        tmax[3 * id + 0] += tma.x; tmax[3 * id + 1] += tma.y; tmax[3 * id + 2] += tma.z;
        rres[id] += res;
    }
    return overflow;
}

/* ----------------------------------------------------------- cylinders (impl:1111-1460) */
typedef struct { v3 pos, base, top, ab; v4 e; } cyl_axis;

static cyl_axis cylinder_axis(const ref_cylinder_info* cy, const double* bi)
{
    cyl_axis a;
    a.pos = mk3(bi[REF_PX], bi[REF_PY], bi[REF_PZ]);
    v4 e = {bi[REF_E0], bi[REF_E1], bi[REF_E2], bi[REF_E3]};
    a.e = e;
    a.base = add3(a.pos, to_global(ld3(cy->pbase), e));   /* impl:1202 */
    a.top = add3(a.pos, to_global(ld3(cy->ptop), e));     /* impl:1204 */
    a.ab = sub3(a.top, a.base);
    return a;
}

/* particle_cylinder_contact_detection, impl:1189-1288 */
static double cylinder_detection(const ref_cylinder_info* cy, cyl_axis a, v3 p, double r, v3* u, int* inner)
{
    double t = dot3(sub3(p, a.base), a.ab) / dot3(a.ab, a.ab);
    v3 cp = add3(a.base, scl3(t, a.ab));
    double dist = len3(sub3(p, cp));
    if (t >= 0 && t <= 1) {                                /* impl:1210-1229 */
        if (dist == 0) { *inner = 1; return 0; }
        *u = div3(sub3(cp, p), dist);
        if (dist < cy->len_rr[2]) {
            *inner = 1;
            *u = neg3(*u);
            return dist + r - cy->len_rr[1];
        }
        return cy->len_rr[1] + r - dist;
    }
    double thick = cy->thickness;                          /* impl:1233-1241 */
    int one = thick ? 1 : 0;
    if (dist < cy->len_rr[2] + thick && dist > one * cy->len_rr[2]) {
        v3 oc = sub3(a.pos, cp);
        double ocl = len3(oc);
        *u = div3(oc, ocl);
        return cy->len_rr[0] * 0.5 + r - ocl;
    }
    v3 at = to_local(sub3(p, a.top), a.e);                 /* impl:1245-1252 */
    v3 rim = a.top;
    if (fabs(at.y) > cy->len_rr[0]) {
        at = to_local(sub3(p, a.base), a.e);
        rim = a.base;
    }
    double rr = cy->len_rr[2];
    double phi = atan(at.x / at.z);                        /* impl:1253-1270 */
    double lx;
    if (phi < 0 && at.z < 0) lx = rr * sin(-phi);
    else if (phi > 0 && at.x < 0 && at.z < 0) lx = rr * sin(-phi);
    else lx = rr * sin(phi);
    double lz = rr * cos(phi);
    if (at.z < 0 && lz > 0) lz = -lz;
    else if (at.z > 0 && lz < 0) lz = -lz;
    v3 c = add3(rim, to_global(mk3(lx, 0.0, lz), a.e));
    v3 dv = sub3(c, p);
    double d = len3(dv);
    *u = div3(dv, d);
    if (d < r) return r - d;
    return -1.0;
}

/* particle_cylinder_inner_base_or_top_contact_detection, impl:1290-1325 */
static double cylinder_inner_lid(const ref_cylinder_info* cy, cyl_axis a, v3 p, double r, v3* u, v3* cpt)
{
    double t = dot3(sub3(p, a.base), a.ab) / dot3(a.ab, a.ab);
    v3 cp = add3(a.base, scl3(t, a.ab));
    double dist = len3(sub3(p, cp));
    if (cy->empty_part == 1 && t > 0.6) return 0.0;
    if (cy->empty_part == 2 && t < 0.4) return 0.0;
    double gab = -1.0;
    if (dist < cy->len_rr[2]) {
        v3 oc = sub3(a.pos, cp);
        double ocl = len3(oc);
        *u = div3(neg3(oc), ocl);
        *cpt = add3(p, scl3(r, *u));
        gab = r + ocl - cy->len_rr[0] * 0.5;
    }
    return gab;
}

double ref_cylinder_detection(const ref_cylinder_info* cy, const double* dbi, const double p[3], double r,
                              double u_io[3], int* inner_out)
{
    v3 u = ld3(u_io);
    int inner = 0;
    double d = cylinder_detection(cy, cylinder_axis(cy, dbi), ld3(p), r, &u, &inner);
    st3(u_io, u);
    if (inner_out) *inner_out = inner;
    return d;
}

double ref_cylinder_inner_lid(const ref_cylinder_info* cy, const double* dbi, const double p[3], double r,
                              double u_io[3], double cpt_io[3])
{
    v3 u = ld3(u_io), c = ld3(cpt_io);
    double d = cylinder_inner_lid(cy, cylinder_axis(cy, dbi), ld3(p), r, &u, &c);
    st3(u_io, u);
    st3(cpt_io, c);
    return d;
}

#define REF_CYL_SLOTS 3

/* one contact (particle_cylinder_contact_force, impl:1111-1187) or its cohesive attraction
 * (impl:1419-1441, 1444-1458), stored in the slot after *n_new */
static void cylinder_pair(const ref_dem_parameters* p, const ref_quirks* q, const ref_contact_property* cp,
                          const double* bi, v3 ip3, double r, double m, v3 ivel, v3 iom, v3 cpt, v3 unit,
                          double cdist, double coh_s, unsigned int pid, const unsigned int* p_id,
                          const double* p_sd, unsigned int* pair_id, double* tsd, unsigned int sid,
                          unsigned int* n_new, int* overflow, v3* sumF, v3* sumM, v3* tma, double* res, double* bf)
{
    v3 bpos = mk3(bi[REF_PX], bi[REF_PY], bi[REF_PZ]);
    double sd0 = 0.0, sd1 = 0.0;
    int stored = 0;
    if (cdist > 0) {
        v3 po2cp = sub3(cpt, bpos);
        v4 be = {bi[REF_E0], bi[REF_E1], bi[REF_E2], bi[REF_E3]};
        v4 bd = {bi[REF_ED0], bi[REF_ED1], bi[REF_ED2], bi[REF_ED3]};
        v3 om = ang_vel(be, bd);
        for (unsigned int s = 0; s < REF_CYL_SLOTS; s++)
            if (p_id[s] == pid) { sd0 = p_sd[2 * s]; sd1 = p_sd[2 * s + 1]; break; }
        v3 rc = sub3(cpt, ip3);
        v3 dv = sub3(add3(mk3(bi[REF_VX], bi[REF_VY], bi[REF_VZ]), cross3(om, po2cp)), add3(ivel, cross3(iom, rc)));
        fconst c = get_constant(p->contact_model, r, 0.0, m, 0.0, cp);
        v3 Fn = mk3(0, 0, 0), Ft = mk3(0, 0, 0);
        force_law(p->contact_model, q->gpu_force_law, p->dt, c, r, 0, cp, cdist, &sd0, &sd1, dv, unit, &Ft, &Fn);
        v3 mf = add3(Fn, Ft);
        *tma = add3(*tma, cross3(rc, mf));       /* calculate_previous_rolling_resistance, impl:476-483 */
        *res += r * cp->rfric * len3(Fn);
        *sumF = add3(*sumF, mf);
        *sumM = add3(*sumM, cross3(rc, mf));
        bf[0] += -mf.x; bf[1] += -mf.y; bf[2] += -mf.z;
        v3 bm = neg3(cross3(po2cp, mf));
        bf[3] += bm.x; bf[4] += bm.y; bf[5] += bm.z;
        stored = 1;
    } else if (cdist <= 0 && fabs(cdist) < fabs(coh_s)) {
        double f = jkr_separation_force(r, 0, cp->coh);
        double cf = ref_cohesion_force(r, 0, cp, f);
        v3 mf = scl3(-cf, unit);
        *sumF = add3(*sumF, mf);
        bf[0] += -mf.x; bf[1] += -mf.y; bf[2] += -mf.z;
        v3 bm = neg3(cross3(sub3(cpt, bpos), mf));
        bf[3] += bm.x; bf[4] += bm.y; bf[5] += bm.z;
        stored = 1;
    }
    if (!stored) return;
    if (*n_new < REF_CYL_SLOTS) {
        tsd[2 * (sid + *n_new)] = sd0;
        tsd[2 * (sid + *n_new) + 1] = sd1;
        pair_id[sid + *n_new] = pid;
        (*n_new)++;
    } else {
        *overflow = 1;
    }
}

unsigned int ref_cylinder_contact_force(const ref_dem_parameters* p, const ref_quirks* q,
                                        const ref_cylinder_info* cy, const double* dbi,
                                        const ref_contact_property* cp,
                                        double* fx, double* fy, double* fz,
                                        double* mx, double* my, double* mz,
                                        const double* pos, const double* ep, const double* vel, const double* ev,
                                        double* force, double* moment, const double* mass,
                                        double* tmax, double* rres,
                                        unsigned int* pair_count, unsigned int* pair_id, double* tsd,
                                        unsigned int np)
{
    unsigned int overflow = 0;
    cyl_axis a = cylinder_axis(cy, dbi);
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+:overflow)
    for (long ii = 0; ii < (long)np; ii++) {
        unsigned int id = (unsigned int)ii;
        unsigned int sid = id * REF_CYL_SLOTS;
        unsigned int p_id[REF_CYL_SLOTS];
        double p_sd[2 * REF_CYL_SLOTS];
        for (unsigned int s = 0; s < REF_CYL_SLOTS; s++) {       /* impl:1337-1343 */
            p_id[s] = pair_id[sid + s];
            p_sd[2 * s] = tsd[2 * (sid + s)];
            p_sd[2 * s + 1] = tsd[2 * (sid + s) + 1];
        }
        double bf[6] = {0, 0, 0, 0, 0, 0};
        unsigned int old_count = pair_count[id];
        unsigned int k = cy->id;
        double m = mass[id];
        v4 ipos = ld4(pos + 4 * id);
        v3 ip3 = mk3(ipos.x, ipos.y, ipos.z);
        double r = ipos.w;
        v3 ivel = ld3(vel + 3 * id);
        v3 iom = ang_vel(ld4(ep + 4 * id), ld4(ev + 4 * id));
        v3 sumF = mk3(0, 0, 0), sumM = mk3(0, 0, 0), tma = mk3(0, 0, 0);
        double res = 0.0;
        unsigned int n_new = k ? old_count : 0;               /* impl:1361: a later cylinder appends */
        if (n_new > REF_CYL_SLOTS) n_new = REF_CYL_SLOTS;
        int inner = 0, ovf = 0;
        v3 unit = mk3(0, 0, 0);
        double cdist = cylinder_detection(cy, a, ip3, r, &unit, &inner);
        double coh_s = 0;
        if (cp->coh) coh_s = ref_cohesion_limit_depth(r, 0, cp);
        v3 cpt = add3(ip3, scl3(r, unit));
        cylinder_pair(p, q, cp, dbi, ip3, r, m, ivel, iom, cpt, unit, cdist, coh_s, k, p_id, p_sd, pair_id, tsd, sid,
                      &n_new, &ovf, &sumF, &sumM, &tma, &res, bf);
        if (inner) {                                          /* impl:1418-1441 */
            cdist = cylinder_inner_lid(cy, a, ip3, r, &unit, &cpt);
            cylinder_pair(p, q, cp, dbi, ip3, r, m, ivel, iom, cpt, unit, cdist, coh_s, k + 1000u, p_id, p_sd, pair_id,
                          tsd, sid, &n_new, &ovf, &sumF, &sumM, &tma, &res, bf);
        }
        overflow += (unsigned int)ovf;
        if (fx) { fx[id] = bf[0]; fy[id] = bf[1]; fz[id] = bf[2]; mx[id] = bf[3]; my[id] = bf[4]; mz[id] = bf[5]; }
        double* F = force + 3 * id;
        double* M = moment + 3 * id;
        F[0] += sumF.x; F[1] += sumF.y; F[2] += sumF.z;
        M[0] += sumM.x; M[1] += sumM.y; M[2] += sumM.z;
        pair_count[id] = n_new;
        tmax[3 * id + 0] += tma.x; tmax[3 * id + 1] += tma.y; tmax[3 * id + 2] += tma.z;
        rres[id] += res;
    }
    return overflow;
}

/* ---------------------------------------------------------- triangles (impl:1655-1746) */
static int closest_pt_triangle(const ref_triangle_info* t, v3 p, v3* out)
{
    v3 a = ld3(t->P), b = ld3(t->Q), c = ld3(t->R);
    v3 ab = sub3(b, a), ac = sub3(c, a), ap = sub3(p, a), bp = sub3(p, b), cp = sub3(p, c);
    double d1 = dot3(ab, ap), d2 = dot3(ac, ap), d3 = dot3(ab, bp);
    double d4 = dot3(ac, bp), d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    double va = d3 * d6 - d5 * d4;
    double vb = d5 * d2 - d1 * d6;
    double vc = d1 * d4 - d3 * d2;
    v3 cpt = mk3(0, 0, 0);
    int ct = -1;
    /* every test runs; a later match overrides an earlier one (vertex, edge, face) */
    if (d1 <= 0.0 && d2 <= 0.0) { ct = 0; cpt = a; }
    if (d3 >= 0.0 && d4 <= d3) { ct = 0; cpt = b; }
    if (d6 >= 0.0 && d5 <= d6) { ct = 0; cpt = c; }
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) { ct = 1; cpt = add3(a, scl3(d1 / (d1 - d3), ab)); }
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) { ct = 1; cpt = add3(a, scl3(d2 / (d2 - d6), ac)); }
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        ct = 1; cpt = add3(b, scl3((d4 - d3) / ((d4 - d3) + (d5 - d6)), sub3(c, b)));
    }
    if (va >= 0 && vb >= 0 && vc >= 0) {
        double denom = 1.0 / (va + vb + vc);
        v3 v = scl3(vb * denom, ab);
        v3 w = scl3(vc * denom, ac);
        cpt = add3(add3(a, v), w);
        ct = 2;
    }
    *out = cpt;
    return ct;
}

int ref_closest_pt_point_triangle(const ref_triangle_info* t, const double p[3], double out[3])
{
    v3 o;
    int ct = closest_pt_triangle(t, ld3(p), &o);
    st3(out, o);
    return ct;
}

static int check_overlap(v3 p, v3 c)
{
    int b0 = (p.x >= c.x - 1e-9 && p.x <= c.x + 1e-9);
    int b1 = (p.y >= c.y - 1e-9 && p.y <= c.y + 1e-9);
    int b2 = (p.z >= c.z - 1e-9 && p.z <= c.z + 1e-9);
    return b0 && b1 && b2;
}

typedef struct { unsigned int id; v3 cpt; } tri_hit;

/* impl:1748-1835 */
static void tri_contact(const ref_dem_parameters* p, const ref_quirks* q, tri_hit h,
                        const double* dbi, const ref_contact_property* cp, double* bf,
                        v3 ipos, v3 ivel, v3 iom, const unsigned int* p_id, const double* p_sd,
                        unsigned int* pair_id, double* tsd, unsigned int sid, double r, double m,
                        v3* sumF, v3* sumM, double* res, v3* tma, unsigned int* n_new,
                        unsigned int* n_seen, unsigned int capacity)
{
    v3 rp = sub3(h.cpt, ipos);
    double sd0 = 0.0, sd1 = 0.0;
    v3 po2cp = sub3(h.cpt, mk3(dbi[REF_PX], dbi[REF_PY], dbi[REF_PZ]));
    double dist = len3(rp);
    v3 unit = div3(rp, dist);
    double cdist = r - dist;
    double coh_s = 0;
    if (cp->coh) coh_s = ref_cohesion_limit_depth(r, 0, cp);
    if (cdist <= 0 && fabs(cdist) < fabs(coh_s)) {
        double f = jkr_separation_force(r, 0, cp->coh);
        double cf = ref_cohesion_force(r, 0, cp, f);
        *sumF = sub3(*sumF, scl3(cf, unit));
        (*n_seen)++;
        if (*n_new < capacity) {
            tsd[2 * (sid + *n_new)] = sd0; tsd[2 * (sid + *n_new) + 1] = sd1;
            pair_id[sid + *n_new] = h.id; (*n_new)++;
        }
        return;
    }
    v3 Fn = mk3(0, 0, 0), Ft = mk3(0, 0, 0);
    v4 be = {dbi[REF_E0], dbi[REF_E1], dbi[REF_E2], dbi[REF_E3]};
    v4 bd = {dbi[REF_ED0], dbi[REF_ED1], dbi[REF_ED2], dbi[REF_ED3]};
    v3 jom = ang_vel(be, bd);
    for (unsigned int s = 0; s < capacity; s++)
        if (p_id[s] == h.id) { sd0 = p_sd[2 * s]; sd1 = p_sd[2 * s + 1]; break; }
    v3 rc = sub3(h.cpt, ipos);
    v3 dv = sub3(add3(mk3(dbi[REF_VX], dbi[REF_VY], dbi[REF_VZ]), cross3(jom, po2cp)), add3(ivel, cross3(iom, rc)));
    fconst c = get_constant(p->contact_model, r, 0, m, dbi[REF_MM], cp);
    force_law(p->contact_model, q->gpu_force_law, p->dt, c, r, 0, cp, cdist, &sd0, &sd1, dv, unit, &Ft, &Fn);
    v3 mf = add3(Fn, Ft);
    *tma = add3(*tma, cross3(rc, mf));
    *res += r * cp->rfric * len3(Fn);
    *sumF = add3(*sumF, mf);
    *sumM = add3(*sumM, cross3(rc, mf));
    bf[0] += -mf.x; bf[1] += -mf.y; bf[2] += -mf.z;
    v3 mm = neg3(cross3(po2cp, mf));
    bf[3] += mm.x; bf[4] += mm.y; bf[5] += mm.z;
    (*n_seen)++;
    if (*n_new < capacity) {
        tsd[2 * (sid + *n_new)] = sd0; tsd[2 * (sid + *n_new) + 1] = sd1;
        pair_id[sid + *n_new] = h.id; (*n_new)++;
    }
}

#define REF_TRI_LIST 20

unsigned int ref_particle_polygon_collision(const ref_dem_parameters* p, const ref_quirks* q,
                                            const ref_triangle_info* dti, const double* dbi,
                                            double* fx, double* fy, double* fz,
                                            double* mx, double* my, double* mz,
                                            const double* pos, const double* ep,
                                            const double* vel, const double* ev,
                                            double* force, double* moment, const double* mass,
                                            double* tmax, double* rres,
                                            unsigned int* pair_count, unsigned int* pair_id,
                                            double* tsd, const unsigned int* sorted_index,
                                            const unsigned int* cstart, const unsigned int* cend,
                                            const ref_contact_property* cp,
                                            unsigned int np, unsigned int ntriangle,
                                            unsigned int capacity)
{
    unsigned int overflow = 0;
    const unsigned int sk = dti[0].tid;
    if (capacity > REF_MAX_CAPACITY) capacity = REF_MAX_CAPACITY;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads) reduction(+:overflow)
    for (long ii = 0; ii < (long)np; ii++) {
        unsigned int id = (unsigned int)ii;
        unsigned int sid = id * capacity;
        unsigned int p_id[REF_MAX_CAPACITY]; double p_sd[2 * REF_MAX_CAPACITY];
        for (unsigned int s = 0; s < capacity; s++) {
            p_id[s] = pair_id[sid + s];
            p_sd[2 * s] = tsd[2 * (sid + s)];
            p_sd[2 * s + 1] = tsd[2 * (sid + s) + 1];
        }
        double im = mass[id];
        v3 ipos = ld3(pos + 4 * id);
        v3 ivel = ld3(vel + 3 * id);
        v3 iom = ang_vel(ld4(ep + 4 * id), ld4(ev + 4 * id));
        int g[3];
        grid_pos(p, ipos, g);
        double ir = pos[4 * id + 3];
        v3 sumF = mk3(0, 0, 0), sumM = mk3(0, 0, 0);
        unsigned int n_new = 0, n_seen = 0;
        tri_hit ctri[REF_TRI_LIST], cline[REF_TRI_LIST], cpnt[REF_TRI_LIST];
        unsigned int nct = 0, ncl = 0, ncp = 0;
        int list_full = 0;
        double coh_s = 0;
        v3 prev_line = mk3(0, 0, 0), prev_point = mk3(0, 0, 0);
        for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
        for (int x = -1; x <= 1; x++) {
            unsigned int h = grid_hash(p, g[0] + x, g[1] + y, g[2] + z);
            unsigned int start = cstart[h];
            if (start == 0xffffffffu) continue;
            unsigned int end = cend[h];
            for (unsigned int j = start; j < end; j++) {
                unsigned int tk = sorted_index[j];
                if (!(tk >= np + sk && tk < np + sk + ntriangle)) continue;
                tk -= np + sk;
                v3 cpt;
                int t = closest_pt_triangle(&dti[tk], ipos, &cpt);
                double cdist = ir - len3(sub3(ipos, cpt));
                if (cp->coh) coh_s = ref_cohesion_limit_depth(ir, 0, cp);
                int take = 0;
                if (cdist > 0) {
                    if (t == 1) {
                        if (check_overlap(prev_line, cpt)) continue;
                        prev_line = cpt;
                    } else if (t == 0) {
                        if (check_overlap(prev_point, cpt)) continue;
                        prev_point = cpt;
                    }
                    take = 1;
                } else if (cdist <= 0 && fabs(cdist) < fabs(coh_s)) {
                    take = 1;
                }
                if (take) {
                    tri_hit hit = {tk, cpt};
                    if (t == 2) { if (nct < REF_TRI_LIST) ctri[nct++] = hit; else list_full = 1; }
                    else if (t == 1) { if (ncl < REF_TRI_LIST) cline[ncl++] = hit; else list_full = 1; }
                    else if (t == 0) { if (ncp < REF_TRI_LIST) cpnt[ncp++] = hit; else list_full = 1; }
                }
            }
        }
        (void)list_full;
        double res = 0.0;
        v3 tma = mk3(0, 0, 0);
        double bf[6] = {0, 0, 0, 0, 0, 0};
        for (unsigned int k = 0; k < nct; k++)
            tri_contact(p, q, ctri[k], dbi, cp, bf, ipos, ivel, iom, p_id, p_sd, pair_id, tsd, sid,
                        ir, im, &sumF, &sumM, &res, &tma, &n_new, &n_seen, capacity);
        if (!nct)
            for (unsigned int k = 0; k < ncl; k++)
                tri_contact(p, q, cline[k], dbi, cp, bf, ipos, ivel, iom, p_id, p_sd, pair_id, tsd, sid,
                            ir, im, &sumF, &sumM, &res, &tma, &n_new, &n_seen, capacity);
        if (!nct && !ncl)
            for (unsigned int k = 0; k < ncp; k++)
                tri_contact(p, q, cpnt[k], dbi, cp, bf, ipos, ivel, iom, p_id, p_sd, pair_id, tsd, sid,
                            ir, im, &sumF, &sumM, &res, &tma, &n_new, &n_seen, capacity);
        if (fx) { fx[id] = bf[0]; fy[id] = bf[1]; fz[id] = bf[2]; mx[id] = bf[3]; my[id] = bf[4]; mz[id] = bf[5]; }
        double* F = force + 3 * id;
        double* M = moment + 3 * id;
        F[0] += sumF.x; F[1] += sumF.y; F[2] += sumF.z;
        M[0] += sumM.x; M[1] += sumM.y; M[2] += sumM.z;
        if (n_seen > capacity) overflow++;
        pair_count[id] = n_new;
        tmax[3 * id + 0] += tma.x; tmax[3 * id + 1] += tma.y; tmax[3 * id + 2] += tma.z;
        rres[id] += res;
    }
    return overflow;
}

/* ------------------------------------------------------------ rolling (impl:2420-2454) */
void ref_decide_rolling_friction_moment(const ref_dem_parameters* p, const ref_quirks* q,
                                        double* tmax, double* rres, const double* inertia,
                                        const double* ep, const double* ev, double* moment,
                                        unsigned int np)
{
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long ii = 0; ii < (long)np; ii++) {
        unsigned int id = (unsigned int)ii;
        double Tr = rres[id];
        if (!Tr) continue;
        v3 Tm = ld3(tmax + 3 * id);
        double J = q->flat_inertia ? inertia[id] : inertia[3 * id];
        v3 w = ang_vel(ld4(ep + 4 * id), ld4(ev + 4 * id));
        double Jdt = J * p->dt;
        v3 T = mk3(Jdt * w.x - Tm.x, Jdt * w.y - Tm.y, Jdt * w.z - Tm.z);
        double lT = len3(T);
        if (lT && Tr) {
            v3 Trv = scl3(Tr, div3(T, lT));
            if (len3(Trv) >= len3(T)) Trv = T;
            moment[3 * id + 0] += Trv.x;
            moment[3 * id + 1] += Trv.y;
            moment[3 * id + 2] += Trv.z;
        }
        tmax[3 * id + 0] = 0.0; tmax[3 * id + 1] = 0.0; tmax[3 * id + 2] = 0.0;
        rres[id] = 0.0;
    }
}

/* ------------------------------------------------------------ mesh pose (impl:2544-2601) */
void ref_update_mesh_object_data(const double* dbi, const double* vList, double* sphere,
                                 const double* dlocal, ref_triangle_info* dpi,
                                 unsigned int ntriangle)
{
    v3 bp = mk3(dbi[REF_PX], dbi[REF_PY], dbi[REF_PZ]);
    v4 e = {dbi[REF_E0], dbi[REF_E1], dbi[REF_E2], dbi[REF_E3]};
    for (unsigned int i = 0; i < ntriangle; i++) {
        const double* s = vList + 9 * i;
        v3 P = add3(bp, to_global(ld3(s + 0), e));
        v3 Q = add3(bp, to_global(ld3(s + 3), e));
        v3 R = add3(bp, to_global(ld3(s + 6), e));
        v3 c = add3(bp, to_global(ld3(dlocal + 3 * i), e));
        sphere[4 * i + 0] = c.x; sphere[4 * i + 1] = c.y; sphere[4 * i + 2] = c.z;
        st3(dpi[i].P, P); st3(dpi[i].Q, Q); st3(dpi[i].R, R);
    }
}

/* ------------------------ mesh definition (xParticleMeshObjectContact.cpp:104-127) */
/* Proxy sphere of every triangle: the triangle placed at the body pose (p = pos + A(e) v),
 * FitSphereToTriangle(p, q, r, 0.8), its centre kept in the body frame (A(e)^T (c - pos)). */
void ref_mesh_define(const double* dbi, const double* vList, unsigned int ntriangle, double* dlocal,
                     double* radius)
{
    v3 bp = mk3(dbi[REF_PX], dbi[REF_PY], dbi[REF_PZ]);
    v4 e = {dbi[REF_E0], dbi[REF_E1], dbi[REF_E2], dbi[REF_E3]};
    for (unsigned int i = 0; i < ntriangle; i++) {
        const double* s = vList + 9 * i;
        double P[3], Q[3], R[3], c[4];
        st3(P, add3(bp, to_global(ld3(s + 0), e)));
        st3(Q, add3(bp, to_global(ld3(s + 3), e)));
        st3(R, add3(bp, to_global(ld3(s + 6), e)));
        ref_fit_sphere_to_triangle(P, Q, R, 0.8, c);
        st3(dlocal + 3 * i, to_local(sub3(mk3(c[0], c[1], c[2]), bp), e));
        radius[i] = c[3];
    }
}

/* ------------------------------------------- wall pose (xParticlePlaneContact.cpp:96-110) */
/* Local corners of a wall at its pose (xPlaneObject::define, xPlaneObject.cpp:45-48). */
void ref_plane_local_points(const ref_plane_info* pl, const double* dbi, double out[12])
{
    v3 bp = mk3(dbi[REF_PX], dbi[REF_PY], dbi[REF_PZ]);
    v4 e = {dbi[REF_E0], dbi[REF_E1], dbi[REF_E2], dbi[REF_E3]};
    const double* c[4] = {pl->xw, pl->w2, pl->w3, pl->w4};
    for (int q = 0; q < 4; q++) st3(out + 3 * q, to_local(sub3(ld3(c[q]), bp), e));
}

/* Corners, edge vectors, lengths and unit vectors of the wall at the pose dbi. */
void ref_plane_update(const double local[12], const double* dbi, unsigned int id, ref_plane_info* out)
{
    v3 bp = mk3(dbi[REF_PX], dbi[REF_PY], dbi[REF_PZ]);
    v4 e = {dbi[REF_E0], dbi[REF_E1], dbi[REF_E2], dbi[REF_E3]};
    v3 xw = add3(bp, to_global(ld3(local + 0), e));
    v3 w2 = add3(bp, to_global(ld3(local + 3), e));
    v3 w3 = add3(bp, to_global(ld3(local + 6), e));
    v3 w4 = add3(bp, to_global(ld3(local + 9), e));
    v3 pa = sub3(w2, xw), pb = sub3(w4, xw);
    double l1 = sqrt(dot3(pa, pa)), l2 = sqrt(dot3(pb, pb));
    v3 u1 = div3(pa, l1), u2 = div3(pb, l2);
    out->id = id;
    out->l1 = l1; out->l2 = l2;
    st3(out->xw, xw); st3(out->w2, w2); st3(out->w3, w3); st3(out->w4, w4);
    st3(out->pa, pa); st3(out->pb, pb);
    st3(out->u1, u1); st3(out->u2, u2); st3(out->uw, cross3(u1, u2));
}

double ref_reduction(const double* in, unsigned int n)
{
    double s = 0.0;
    for (unsigned int i = 0; i < n; i++) s += in[i];
    return s;
}

/* ---------------------------------------- triangle proxy sphere (xUtilityFunction.cpp:379-415) */
void ref_fit_sphere_to_triangle(const double Pp[3], const double Qp[3], const double Rp[3],
                                double ft, double out[4])
{
    v3 P = ld3(Pp), Q = ld3(Qp), R = ld3(Rp);
    v3 V = sub3(Q, P), W = sub3(R, P);
    v3 N = cross3(V, W);
    N = div3(N, len3(N));
    v3 M1 = div3(add3(Q, P), 2.0), M2 = div3(add3(R, P), 2.0);
    v3 D1 = cross3(N, V), D2 = cross3(N, W);
    double t = 0.0;
    if (fabs(D1.x * D2.y - D1.y * D2.x) > 1E-13)
        t = (D2.x * (M1.y - M2.y)) / (D1.x * D2.y - D1.y * D2.x) - (D2.y * (M1.x - M2.x)) / (D1.x * D2.y - D1.y * D2.x);
    else if (fabs(D1.x * D2.z - D1.z * D2.x) > 1E-13)
        t = (D2.x * (M1.z - M2.z)) / (D1.x * D2.z - D1.z * D2.x) - (D2.z * (M1.x - M2.x)) / (D1.x * D2.z - D1.z * D2.x);
    else if (fabs(D1.y * D2.z - D1.z * D2.y) > 1E-13)
        t = (D2.y * (M1.z - M2.z)) / (D1.y * D2.z - D1.z * D2.y) - (D2.z * (M1.y - M2.y)) / (D1.y * D2.z - D1.z * D2.y);
    v3 Ctri = add3(M1, scl3(t, D1));
    v3 Cs = mk3(0, 0, 0);
    double fc = 0, tol = 1e-4;
    double r = len3(sub3(P, Ctri));
    int guard = 0;
    while (fabs(fc - ft) > tol && guard++ < 1000) {
        double d = ft * r;
        double pp = d / len3(N);
        Cs = sub3(Ctri, scl3(pp, N));
        r = len3(sub3(P, Cs));
        fc = d / r;
    }
    out[0] = Cs.x; out[1] = Cs.y; out[2] = Cs.z; out[3] = r;
}

/* ------------------------------------------------- step driver (xIntegratorVV.cpp:26-64) */
int ref_one_step(ref_sim* s)
{
    const unsigned int np = s->np;
    ref_dem_parameters* p = &s->prm;
    unsigned int ntri = 0;
    for (unsigned int m = 0; m < s->nmesh; m++) ntri += s->meshes[m].ntriangle;
    const unsigned int n = np + ntri;
    const unsigned int ncell = p->grid_size[0] * p->grid_size[1] * p->grid_size[2];

    ref_vv_update_position(p, s->pos, s->ep, s->vel, s->ev, s->acc, s->ea, np);
    ref_calculate_hash_and_index(p, s->hash, s->index, s->pos, 0, np);
    unsigned int sid = np;
    for (unsigned int m = 0; m < s->nmesh; m++) {
        ref_calculate_hash_and_index(p, s->hash, s->index, s->meshes[m].sphere, sid, s->meshes[m].ntriangle);
        sid += s->meshes[m].ntriangle;
    }
    ref_reorder_data_and_find_cell_start(s->hash, s->index, s->cstart, s->cend, s->sorted_index, n, ncell);

    if (s->has_pp)
        s->overflow_pp += ref_calculate_p2p(p, &s->q, s->pos, s->ep, s->vel, s->ev, s->force, s->moment,
                                            s->mass, s->tmax, s->rres, s->pp_count, s->pp_id, s->pp_tsd,
                                            s->sorted_index, s->cstart, s->cend, &s->cp_pp, np, s->cap_pp);
    for (unsigned int k = 0; k < s->nplane; k++) {
        s->overflow_pl += ref_plane_contact_force(p, &s->q, &s->planes[k], s->plane_dbi + 15 * k, &s->plane_cp[k],
                                                  s->fx, s->fy, s->fz, s->mx, s->my, s->mz,
                                                  s->pos, s->ep, s->vel, s->ev, s->force, s->moment,
                                                  s->mass, s->tmax, s->rres, s->pl_count, s->pl_id, s->pl_tsd,
                                                  np, s->cap_pl);
        if (s->plane_wrench) {
            s->plane_wrench[6 * k + 0] = ref_reduction(s->fx, np);
            s->plane_wrench[6 * k + 1] = ref_reduction(s->fy, np);
            s->plane_wrench[6 * k + 2] = ref_reduction(s->fz, np);
            s->plane_wrench[6 * k + 3] = ref_reduction(s->mx, np);
            s->plane_wrench[6 * k + 4] = ref_reduction(s->my, np);
            s->plane_wrench[6 * k + 5] = ref_reduction(s->mz, np);
        }
    }
    for (unsigned int m = 0; m < s->nmesh; m++) {
        ref_mesh* me = &s->meshes[m];
        s->overflow_ms += ref_particle_polygon_collision(p, &s->q, me->dti, me->dbi,
                                                         s->fx, s->fy, s->fz, s->mx, s->my, s->mz,
                                                         s->pos, s->ep, s->vel, s->ev, s->force, s->moment,
                                                         s->mass, s->tmax, s->rres, s->ms_count, s->ms_id,
                                                         s->ms_tsd, s->sorted_index, s->cstart, s->cend,
                                                         &me->cp, np, me->ntriangle, s->cap_ms);
        me->wrench[0] = ref_reduction(s->fx, np);
        me->wrench[1] = ref_reduction(s->fy, np);
        me->wrench[2] = ref_reduction(s->fz, np);
        me->wrench[3] = ref_reduction(s->mx, np);
        me->wrench[4] = ref_reduction(s->my, np);
        me->wrench[5] = ref_reduction(s->mz, np);
    }
    ref_decide_rolling_friction_moment(p, &s->q, s->tmax, s->rres, s->inertia, s->ep, s->ev, s->moment, np);
    ref_vv_update_velocity(p, &s->q, s->vel, s->acc, s->ep, s->ev, s->ea, s->force, s->moment,
                           s->mass, s->inertia, np);
    return 0;
}

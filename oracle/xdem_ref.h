/*
 * xdem_ref.h -- ORACLE (test infrastructure only).
 *
 * A plain-C, FP64 restatement of the reference DEM time step, used to check the
 * HIP implementation. Nothing in the product links, loads or calls this code; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and
 * only as the checker or as the timed CPU baseline.
 *
 * Every function names the reference lines it restates. The reference cannot be
 * built in this image (Windows headers, CUDA toolkit, LAPACKE), so there is no
 * compiled reference beside this restatement; its functions are pinned instead by the
 * tests in tests/test_oracle_*.py (closed forms, brute force, golden fixtures).
 *
 * Deliberate restatement choices (each keeps the reference's mathematics):
 *   - pow(x, 1.5) is evaluated as x*sqrt(x) and pow(x, 2.0) as x*x, so that the CPU
 *     and GPU agree bit-for-bit without depending on two different libm pow's;
 *   - history overflow (more contacts than slots) never writes past the slots: the
 *     contact still acts, its history is not stored, and the overflow is counted
 *     (reference: silent write into the next particle's slots, impl:704-706).
 *
 * Array layouts are the reference's AoS layouts (pos double4, vel double3, ...).
 */
#ifndef XDEM_REF_H
#define XDEM_REF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same bytes as device_dem_parameters (decl:128-150). */
typedef struct ref_dem_parameters {
    unsigned char rollingCondition;
    unsigned char pad_[3];
    int contact_model;
    unsigned int np, nmp, nCluster, nClusterObject;
    unsigned int nTsdaConnection, nTsdaConnectionList, nTsdaConnectionBodyData;
    unsigned int nsphere, nplane, ncylinder, ncell;
    unsigned int grid_size[3];
    double dt, half2dt, cell_size, cohesion;
    double gravity[3];
    double world_origin[3];
} ref_dem_parameters;

/* Same bytes as device_contact_property (decl:230-243). */
typedef struct ref_contact_property {
    double Ei, Ej, pri, prj, Gi, Gj, rest, fric, s_fric, rfric, coh, sratio, amp;
} ref_contact_property;

/* Same bytes as device_plane_info (decl:161-174). */
typedef struct ref_plane_info {
    unsigned int id, pad_;
    double l1, l2;
    double u1[3], u2[3], uw[3], xw[3], pa[3], pb[3], w2[3], w3[3], w4[3];
} ref_plane_info;

/* Same bytes as device_triangle_info (decl:152-159). */
typedef struct ref_triangle_info {
    int id;
    unsigned int tid;
    double P[3], Q[3], R[3];
} ref_triangle_info;

/* Quirk switches: behaviours of the reference that are kept or changed (defaults give the
 * parity profile). */
/* device_cylinder_info (decl:220-228): len_rr = (length, side-wall radius, inner/rim radius). */
typedef struct ref_cylinder_info {
    unsigned int id, empty_part;
    double thickness;
    double len_rr[3];
    double pbase[3], ptop[3];
} ref_cylinder_info;

typedef struct ref_quirks {
    int gpu_force_law;        /* Q2/Q3: 0 = CPU formula (xContact.cpp:238-264), 1 = kernel text (impl:432-452) */
    int velocity_not_updated; /* Q1: 1 = kernel text stores the old velocity (impl:207-221); 0 = CPU Verlet */
    int stale_ft;             /* Q4: 1 = Ft carries over from the previous pair when |e| = 0 (impl:645) */
    int flat_inertia;         /* Q5: 1 = rolling reads ((double*)inertia)[id] (impl:2437) */
} ref_quirks;

/* Byte offsets of the body pose doubles (decl:20). */
enum { REF_MM = 0, REF_PX, REF_PY, REF_PZ, REF_VX, REF_VY, REF_VZ,
       REF_E0, REF_E1, REF_E2, REF_E3, REF_ED0, REF_ED1, REF_ED2, REF_ED3 };

void ref_set_threads(int n);
int ref_get_threads(void);

/* impl:135-147 */
void ref_vv_update_position(const ref_dem_parameters* p, double* pos, double* ep,
                            const double* vel, const double* ev, const double* acc,
                            const double* ea, unsigned int np);

/* impl:192-222 (velocity_not_updated = 1) or xIntegratorVV.cpp:131-152 (= 0);
 * the Euler-parameter acceleration always follows calculate_uceom, impl:74-133. */
void ref_vv_update_velocity(const ref_dem_parameters* p, const ref_quirks* q,
                            double* vel, double* acc, const double* ep, double* ev,
                            double* ea, double* force, double* moment,
                            const double* mass, const double* iner, unsigned int np);

/* impl:54-72, 290-302 */
void ref_calculate_hash_and_index(const ref_dem_parameters* p, unsigned int* hash,
                                  unsigned int* index, const double* pos,
                                  unsigned int sid, unsigned int np);

/* def:94-124 + impl:317-360: stable sort of (hash, index), full clear of the cell
 * tables, cell bounds, sorted index. hash and index are sorted in place. */
void ref_reorder_data_and_find_cell_start(unsigned int* hash, unsigned int* index,
                                          unsigned int* cstart, unsigned int* cend,
                                          unsigned int* sorted_index, unsigned int n,
                                          unsigned int ncell);

/* When set (non-NULL), ref_calculate_p2p writes out[id] = sum over the sphere's contacts of
 * |F_pair| (the normaliser of the mixed-precision error metric). */
void ref_set_pair_magnitude_output(double* out);

/* impl:613-733. Returns the number of particles whose contacts exceeded `capacity`. */
unsigned int ref_calculate_p2p(const ref_dem_parameters* p, const ref_quirks* q,
                               const double* pos, const double* ep, const double* vel,
                               const double* ev, double* force, double* moment,
                               const double* mass, double* tmax, double* rres,
                               unsigned int* pair_count, unsigned int* pair_id, double* tsd,
                               const unsigned int* sorted_index, const unsigned int* cstart,
                               const unsigned int* cend, const ref_contact_property* cp,
                               unsigned int np, unsigned int capacity);

/* impl:735-815 + 950-1080, one wall per call. fx..mz may be NULL. */
unsigned int ref_plane_contact_force(const ref_dem_parameters* p, const ref_quirks* q,
                                     const ref_plane_info* plane, const double* dbi,
                                     const ref_contact_property* cp,
                                     double* fx, double* fy, double* fz,
                                     double* mx, double* my, double* mz,
                                     const double* pos, const double* ep, const double* vel,
                                     const double* ev, double* force, double* moment,
                                     const double* mass, double* tmax, double* rres,
                                     unsigned int* pair_count, unsigned int* pair_id,
                                     double* tsd, unsigned int np, unsigned int capacity);

/* impl:1655-1705, 1729-1835, 1930-2109, one mesh per call. fx..mz may be NULL. */
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
                                            unsigned int capacity);

/* impl:2420-2454 */
void ref_decide_rolling_friction_moment(const ref_dem_parameters* p, const ref_quirks* q,
                                        double* tmax, double* rres, const double* inertia,
                                        const double* ep, const double* ev, double* moment,
                                        unsigned int np);

/* impl:2544-2601 */
void ref_update_mesh_object_data(const double* dbi, const double* vList, double* sphere,
                                 const double* dlocal, ref_triangle_info* dpi,
                                 unsigned int ntriangle);

/* xParticleMeshObjectContact::define (xParticleMeshObjectContact.cpp:104-127): proxy spheres
 * fitted to the triangles at pose dbi, centres returned in the body frame. */
void ref_mesh_define(const double* dbi, const double* vList, unsigned int ntriangle, double* dlocal,
                     double* radius);

/* xPlaneObject::define's local corners A(e)^T (corner - position) (xPlaneObject.cpp:45-48),
 * and xParticlePlaneContact::update (xParticlePlaneContact.cpp:96-110): the wall at pose dbi. */
void ref_plane_local_points(const ref_plane_info* pl, const double* dbi, double out[12]);
void ref_plane_update(const double local[12], const double* dbi, unsigned int id, ref_plane_info* out);

/* def:285-317: sum of n doubles, left to right. */
double ref_reduction(const double* in, unsigned int n);

/* Single-function restatements exposed for unit tests. */
int ref_closest_pt_point_triangle(const ref_triangle_info* t, const double p[3], double out[3]);
double ref_plane_detection(const ref_plane_info* pe, const double xp[3], double r, double u_out[3]);
/* Cylinder regions (impl:1189-1325): the outer/inner wall, lid ring and rim test, and the lid
 * seen from inside; inner_out receives the inner flag, cpt_io the contact point (in/out). */
double ref_cylinder_detection(const ref_cylinder_info* cy, const double* dbi, const double p[3], double r,
                              double u_io[3], int* inner_out);
double ref_cylinder_inner_lid(const ref_cylinder_info* cy, const double* dbi, const double p[3], double r,
                              double u_io[3], double cpt_io[3]);
/* cylinder_contact_force_kernel (impl:1327-1460) over np spheres, 3 slots per sphere;
 * returns the number of spheres with more entries than slots (entries past them dropped). */
unsigned int ref_cylinder_contact_force(const ref_dem_parameters* p, const ref_quirks* q,
                                        const ref_cylinder_info* cy, const double* dbi,
                                        const ref_contact_property* cp,
                                        double* fx, double* fy, double* fz,
                                        double* mx, double* my, double* mz,
                                        const double* pos, const double* ep, const double* vel, const double* ev,
                                        double* force, double* moment, const double* mass,
                                        double* tmax, double* rres,
                                        unsigned int* pair_count, unsigned int* pair_id, double* tsd,
                                        unsigned int np);
void ref_calculate_uceom(const double J[3], const double ep[4], const double ev[4],
                         const double n_prime[3], double out[4]);
void ref_to_angular_velocity(const double e[4], const double d[4], double out[3]);
/* One pair: force law as restated (model from p->contact_model, law from q).
 * io_sd holds (ds, s_dot_prev) in and out; io_ft holds the carried Ft in and out. */
void ref_pair_force(const ref_dem_parameters* p, const ref_quirks* q,
                    const ref_contact_property* cp, double ir, double jr, double im, double jm,
                    double cdist, const double dv[3], const double unit[3],
                    double io_sd[2], double io_ft[3], double fn_out[3]);
double ref_cohesion_limit_depth(double ir, double jr, const ref_contact_property* cp);
double ref_cohesion_force(double ir, double jr, const ref_contact_property* cp, double fn);

/* xUtilityFunction.cpp:379-415 (host geometry used at mesh definition). */
void ref_fit_sphere_to_triangle(const double P[3], const double Q[3], const double R[3],
                                double ft, double out[4]);

/* ---------------------------------------------------------------------------------
 * Whole-step driver: xIntegratorVV::OneStepSimulation (xIntegratorVV.cpp:26-64) with
 * xContactManager::deviceCollision (xContactManager.cpp:458-472). Contacts run in the
 * order particle-particle, walls (by id), meshes (by index), then rolling resistance.
 * --------------------------------------------------------------------------------- */
typedef struct ref_mesh {
    unsigned int ntriangle;
    ref_triangle_info* dti;   /* [ntriangle] */
    double* sphere;           /* [4*ntriangle] proxy spheres (x,y,z,r) */
    double* vlist;            /* [9*ntriangle] local vertices */
    double* dlocal;           /* [3*ntriangle] local proxy centres */
    double dbi[15];
    ref_contact_property cp;
    double wrench[6];         /* out: -sum of particle forces and moments on this body */
} ref_mesh;

typedef struct ref_sim {
    ref_dem_parameters prm;
    ref_quirks q;
    unsigned int np;
    double *pos, *vel, *acc, *ep, *ev, *ea, *force, *moment, *mass, *inertia;
    double *tmax, *rres;
    /* sphere contact history */
    ref_contact_property cp_pp;
    int has_pp;
    unsigned int cap_pp;
    unsigned int *pp_count, *pp_id; double* pp_tsd;
    /* walls */
    unsigned int nplane;
    ref_plane_info* planes;
    double* plane_dbi;           /* [15*nplane] */
    ref_contact_property* plane_cp; /* [nplane] */
    double* plane_wrench;        /* out [6*nplane] */
    unsigned int cap_pl;
    unsigned int *pl_count, *pl_id; double* pl_tsd;
    /* meshes (share one history, as the reference's statics do) */
    unsigned int nmesh;
    ref_mesh* meshes;
    unsigned int cap_ms;
    unsigned int *ms_count, *ms_id; double* ms_tsd;
    /* grid work arrays: n = np + sum of triangles; ncell = product of grid_size */
    unsigned int *hash, *index, *sorted_index, *cstart, *cend;
    /* body-force work arrays [np] */
    double *fx, *fy, *fz, *mx, *my, *mz;
    /* overflow counters accumulated over steps */
    unsigned long long overflow_pp, overflow_pl, overflow_ms;
} ref_sim;

int ref_one_step(ref_sim* s);

#ifdef __cplusplus
}
#endif
#endif

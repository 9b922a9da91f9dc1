/*
 * xdem_host.h -- host-side set-up helpers of the MI355X DEM step (plain C ABI, no GPU).
 *
 * These restate the host code that the reference runs before its kernels:
 *   xdem_make_plane           xPlaneObject::define (source/xdynamics_object/xPlaneObject.cpp:22-49)
 *                             + xParticlePlaneContact::update (source/xdynamics_object/xParticlePlaneContact.cpp:96-131)
 *   xdem_mass_inertia         xParticleManager::SetMassAndInertia (source/xdynamics_manager/xParticleManager.cpp:1111-1131)
 *   xdem_critical_time_step   xDiscreteElementMethodSimulation::CriticalTimeStep
 *                             (source/xdynamics_simulation/xDiscreteElementMethodSimulation.cpp:301-311)
 *   xdem_contact_property_make xContact::define (source/xdynamics_object/xContact.cpp:304-322)
 *   xdem_fit_sphere_to_triangle xUtilityFunctions::FitSphereToTriangle (source/xdynamics_algebra/xUtilityFunction.cpp:379-415)
 * and add seeded scene generators for the benchmark workloads (the reference reads its
 * scenes from Excel sheets, which this build does not parse).
 */
#ifndef XDEM_HOST_H
#define XDEM_HOST_H

#include "xdem_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A scene: particles, walls, materials and step settings, all in host memory. */
typedef struct xdem_scene {
    unsigned int np;
    double* pos;       /* [4*np] x, y, z, radius */
    double* vel;       /* [3*np] */
    double* mass;      /* [np] */
    double* inertia;   /* [3*np] */
    unsigned int nplane;
    xdem_plane_info* planes;          /* [nplane], ids 0..nplane-1 */
    double* plane_dbi;                /* [15*nplane] body pose, MM..ED3 */
    xdem_contact_property* plane_cp;  /* [nplane] */
    xdem_contact_property cp_pp;      /* sphere-sphere material pair */
    int contact_model;                /* 1 = Hertz-Mindlin */
    double dt;
    unsigned int nsteps;
    double gravity[3];
    double world_origin[3];
    double cell_size;                 /* 2 x largest radius */
    double density;
    double min_radius, max_radius;
} xdem_scene;

/* Workload ids for xdem_scene_build. */
#define XDEM_SCENE_BOX_LATTICE   1  /* 10,000 spheres r=1 mm on a 2.2 mm jittered lattice, 0.1x0.1x0.3 m box */
#define XDEM_SCENE_DAM_BREAK     2  /* 1,000,000 spheres r~U[1,2] mm, 0.3x0.3x0.6 m column, 1.2x0.3x0.7 m box */
#define XDEM_SCENE_WHEEL_BED     3  /* 500,000 spheres r~U[2,3] mm over a 1.0x0.4 m bin floor, loose (solid fraction
                                     ~0.3, twice the 0.15 m bed height), to settle before the wheel runs; the
                                     wheel is a mesh body (xdem_wheel_mesh) */
#define XDEM_SCENE_BATCH_MEMBER  4  /* one box-settling scene of the 256-scene sweep (index = variant);
                                     dt = min(1e-5 s, 0.2 x Rayleigh time of its smallest sphere) */
#define XDEM_SCENE_PILE          5  /* 20,000 spheres r~U[1,1.5] mm in a 30 mm cylinder region over a floor */

/* Build a scene. np_override > 0 replaces the particle count (same generator, same
 * density); variant selects the member of a batched sweep (ignored otherwise).
 * Returns XDEM_OK or an error; free with xdem_scene_free. */
int xdem_scene_build(int workload, unsigned long long seed, unsigned int np_override,
                     unsigned int variant, xdem_scene* out);
void xdem_scene_free(xdem_scene* s);

/* Triangles of a grousered wheel in its body frame (axis along y, centred at the origin): a
 * rim of n_theta x n_y quads, two side rings (annuli of depth grouser_height, n_ring quads
 * each), and n_grouser bars of the given height and thickness across the rim (top, two
 * sides of n_y quads, two end caps). 2 n_theta n_y + 4 n_ring + n_grouser (6 n_y + 4)
 * triangles: 4,096 for (170, 8, 12 grousers, 188). With vertices NULL, only the count is
 * returned; else *ntriangle holds the capacity on entry. */
int xdem_wheel_mesh(double radius, double width, unsigned int n_theta, unsigned int n_y, unsigned int n_grouser,
                    double grouser_height, double grouser_thickness, unsigned int n_ring, double* vertices,
                    unsigned int* ntriangle);

/* Reference wall set-up: four corners p0..p3 (p0 -> p1 is u1, p0 -> p3 is u2, the
 * normal is u1 x u2), identity orientation, zero velocity. */
int xdem_make_plane(const double p0[3], const double p1[3], const double p2[3],
                    const double p3[3], unsigned int id, double mass,
                    xdem_plane_info* out, double dbi[15]);

void xdem_mass_inertia(const double* pos, unsigned int n, double density,
                       double* mass, double* inertia);

double xdem_critical_time_step(double min_radius, double density, double youngs, double poisson);

void xdem_contact_property_make(double Ei, double Ej, double nui, double nuj, double rest,
                                double fric, double rfric, double coh, double sratio,
                                xdem_contact_property* out);

void xdem_fit_sphere_to_triangle(const double P[3], const double Q[3], const double R[3],
                                 double ft, double out[4]);

/* ---------------------------------------------------------------- reference generators
 * The reference's particle generators (xParticleManager.cpp:646-880) with its data structs
 * (include/xTypes.h:244-261), its layouts and its draws from frand() = rand() / RAND_MAX. rand()
 * is the C library's of the reference's platform (Windows): x <- 214013 x + 2531011 (mod 2^32),
 * result bits 16..30, RAND_MAX 32767, initial state 1; here it runs on the caller's state.
 * With pos == NULL each returns the particle count in *np; otherwise *np is the capacity of pos
 * ([4 * np] x, y, z, r) and receives the count. Sizes under two sphere diameters per counted
 * axis (which wrap the reference's unsigned counts) are refused. */
typedef struct xdem_cube_particle_data { double dx, dy, dz, lx, ly, lz, minr, maxr; } xdem_cube_particle_data;
typedef struct xdem_plane_particle_data {
    double dx, dy, dz, drx, dry, drz, lx, ly, lz, minr, maxr;
} xdem_plane_particle_data;
typedef struct xdem_circle_particle_data {
    double diameter, sx, sy, sz, dx, dy, dz, minr, maxr;
} xdem_circle_particle_data;
unsigned int xdem_crt_rand(unsigned int* state);
double xdem_crt_frand(unsigned int* state);
/* CreateCubeParticle (xParticleManager.cpp:646-712): a lattice of (int)(d / 2 maxr) - 1 per axis,
 * spacing 2 maxr + space, each coordinate jittered by maxr * space * frand() (z drawn first),
 * then radii minr + (maxr - minr) frand() when they differ. */
int xdem_create_cube_particles(const xdem_cube_particle_data* d, unsigned int* rand_state, double* pos,
                               unsigned int* np);
/* CreatePlaneParticle (xParticleManager.cpp:714-792): a layer of nx x nz spheres in the plane of
 * pu, qu through (lx, ly, lz), repeated ceil(dy) times at the same place (dy counts layers).
 * reference_quirks = 1 also reproduces its copy loop with equal radii: every sphere written to
 * pos[0] (the last stays) and the others zero. */
int xdem_create_plane_particles(const xdem_plane_particle_data* d, unsigned int* rand_state, int reference_quirks,
                                double* pos, unsigned int* np);
/* CreateCircleParticle (xParticleManager.cpp:799-880): rings i = 1..nr of radius i (2 maxr + space)
 * about (sx, sy, sz) in the plane of pu, qu, floor(2 pi / dth) spheres each; radii drawn from the
 * caller's state (the reference seeds with the clock). */
int xdem_create_circle_particles(const xdem_circle_particle_data* d, unsigned int* rand_state, double* pos,
                                 unsigned int* np);

/* Deterministic 64-bit generator used by the scene builders (SplitMix64). */
unsigned long long xdem_rng_next(unsigned long long* state);
double xdem_rng_uniform(unsigned long long* state);

#ifdef __cplusplus
}
#endif
#endif

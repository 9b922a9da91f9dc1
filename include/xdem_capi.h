/*
 * xdem_capi.h -- C ABI of the MI355X-native DEM time step (libxdem.so).
 *
 * Two layers, both `extern "C"`, plain pointers and sizes, every entry returns an int
 * status (XDEM_OK = 0, negative on error; see xdem_types.h):
 *
 * 1. Launcher layer. One entry per launcher of the reference's CUDA layer
 *    (include/xdynamics_parallel/xParallelDEM_decl.cuh:280-424, "decl" below), same
 *    argument order and meaning, device pointers in the reference's AoS layouts, queued
 *    on the stream set by xdem_set_stream (default stream otherwise). The C++ header
 *    xParallelDEM_decl.hpp re-exports them under the reference's exact names and
 *    signatures.
 *
 * 2. Step layer (xdem_ctx). The whole DEM step of
 *    xIntegratorVV::OneStepSimulation (source/xdynamics_simulation/xIntegratorVV.cpp:26-64)
 *    with all contact objects (xContactManager::deviceCollision,
 *    source/xdynamics_manager/xContactManager.cpp:458-472) fused into a few kernels, for
 *    one or many independent scenes advanced together. The C++ facade
 *    xDiscreteElementMethodSimulation.hpp is built on it.
 */
#ifndef XDEM_CAPI_H
#define XDEM_CAPI_H

#include "xdem_host.h"
#include "xdem_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ====================================================================== launcher layer */

/* setDEMSymbolicParameter (decl:280; def:16-19). Copies the struct; later launches use it. */
int xdem_setDEMSymbolicParameter(const xdem_dem_parameters* h_paras);
/* Quirk switches for the launchers (default: parity profile, see xdem_quirks). */
int xdem_set_quirks(const xdem_quirks* q);
/* hipStream_t for all launcher-layer work (NULL = default stream). */
int xdem_set_stream(void* hip_stream);
/* Per-particle body-force buffers written by the wall and mesh launchers; the reference
 * keeps them as statics of xContact (xContact::deviceBodyForceX..MomentZ,
 * source/xdynamics_object/xContact.cpp:339-345). NULLs select library-owned buffers. */
int xdem_set_body_force_buffers(double* fx, double* fy, double* fz, double* mx, double* my, double* mz);
int xdem_get_body_force_buffers(double** fx, double** fy, double** fz, double** mx, double** my, double** mz);
/* Device error word of the launcher layer (XDEM_ERR_* bits); clear != 0 resets it. */
int xdem_get_error_word(unsigned int* out, int clear);

/* Launcher workspace (no counterpart in the reference, whose launchers take thrust
 * temporaries, def:107-109, def:290). The launchers never allocate: the error word, the
 * sort buffers of xdem_cu_reorderDataAndFindCellStart, the partials of xdem_reduction and the
 * body-force buffers used when none are registered live in a workspace for launches of up to
 * n elements. xdem_setDEMSymbolicParameter reserves one for np + nsphere elements (an
 * init-time allocation); xdem_reserve_workspace grows the library's for larger launches, and
 * xdem_set_workspace hands the library a caller-owned device buffer instead (256-byte aligned,
 * at least xdem_workspace_size(n) bytes; xdem_set_workspace(NULL, 0, 0) detaches it, after
 * which the caller may free it; the next init call reserves a library workspace). A launch
 * beyond the workspace returns XDEM_E_INVALID (XDEM_E_NOT_READY without one) and sets
 * XDEM_ERR_WORKSPACE in the error word. */
int xdem_workspace_size(unsigned int n, unsigned long long* bytes);
int xdem_reserve_workspace(unsigned int n);
int xdem_set_workspace(void* d_ptr, unsigned long long bytes, unsigned int n);

/* vv_update_position (decl:282-283) */
int xdem_vv_update_position(double* pos, double* ep, double* vel, double* ev, double* acc, double* ea,
                            unsigned int np);
/* vv_update_velocity (decl:284-285) */
int xdem_vv_update_velocity(double* vel, double* acc, double* ep, double* ev, double* ea, double* force,
                            double* moment, double* mass, double* iner, unsigned int np);
/* cu_calculateHashAndIndex (decl:287) */
int xdem_cu_calculateHashAndIndex(unsigned int* hash, unsigned int* index, double* pos, unsigned int sid,
                                  unsigned int np);
/* cu_calculateHashAndIndexForPolygonSphere (decl:288-290) */
int xdem_cu_calculateHashAndIndexForPolygonSphere(unsigned int* hash, unsigned int* index, unsigned int sid,
                                                  unsigned int nsphere, double* sphere);
/* cu_reorderDataAndFindCellStart (decl:291): sorts (hash, index) in place by hash (stable),
 * fills cstart/cend (0xffffffff / 0 for empty cells) and sorted_index. */
int xdem_cu_reorderDataAndFindCellStart(unsigned int* hash, unsigned int* index, unsigned int* cstart,
                                        unsigned int* cend, unsigned int* sorted_index, unsigned int np,
                                        unsigned int ncell);
/* cu_calculate_p2p (decl:293-299); cp is a device pointer, 12 history slots per particle. */
int xdem_cu_calculate_p2p(double* pos, double* ep, double* vel, double* omega, double* force, double* moment,
                          double* mass, double* tmax, double* rres, unsigned int* pair_count,
                          unsigned int* pair_id, double* tsd, unsigned int* sorted_index, unsigned int* cstart,
                          unsigned int* cend, xdem_contact_property* cp, unsigned int np);
/* cu_plane_contact_force (decl:302-308); plan, dbi (15 doubles) and cp are device pointers,
 * 3 wall slots per particle; writes the registered body-force buffers. */
int xdem_cu_plane_contact_force(xdem_plane_info* plan, double* dbi, xdem_contact_property* cp, double* pos,
                                double* ep, double* vel, double* omega, double* force, double* moment,
                                double* mass, double* tmax, double* rres, unsigned int* pair_count,
                                unsigned int* pair_id, double* tsd, unsigned int np);
/* cu_cylinder_contact_force (decl:326-333); cyl, dbi (15 doubles) and cp are device
 * pointers, 3 cylinder slots per particle (ids: the cylinder's id for its side wall, lid ring
 * and rim, id + 1000 for a lid seen from inside); writes the registered body-force buffers.
 * Slots past the third are not written (XDEM_ERR_CYL_OVERFLOW; the reference writes them into
 * the next particle's). */
int xdem_cu_cylinder_contact_force(xdem_cylinder_info* cyl, double* dbi, xdem_contact_property* cp, double* pos,
                                   double* ep, double* vel, double* omega, double* force, double* moment,
                                   double* mass, double* tmax, double* rres, unsigned int* pair_count,
                                   unsigned int* pair_id, double* tsd, unsigned int np);
/* cu_particle_polygonObject_collision (decl:317-324); 5 triangle slots per particle. */
int xdem_cu_particle_polygonObject_collision(xdem_triangle_info* dpi, double* dbi, double* pos, double* ep,
                                             double* vel, double* omega, double* force, double* moment,
                                             double* mass, double* tmax, double* rres, unsigned int* pair_count,
                                             unsigned int* pair_id, double* tsd, double* dsph,
                                             unsigned int* sidx, unsigned int* cstart, unsigned int* cend,
                                             xdem_contact_property* cp, unsigned int np, unsigned int ntriangle);
/* No counterpart in the reference. For callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation): the same sort and cell
 * bounds as xdem_cu_reorderDataAndFindCellStart, but its sort buffers live in a file-scope
 * device array of the library (16.25 MiB per GPU, held until the process ends), so it needs
 * no workspace and calls no allocating routine. At most 1,048,576 elements per call (larger np:
 * XDEM_E_INVALID, nothing queued). Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams or threads) race. */
int xdem_cu_reorderDataAndFindCellStart_noalloc(unsigned int* hash, unsigned int* index, unsigned int* cstart,
                                                unsigned int* cend, unsigned int* sorted_index, unsigned int np,
                                                unsigned int ncell);
/* cu_decide_rolling_friction_moment (decl:336-343) */
int xdem_cu_decide_rolling_friction_moment(double* tmax, double* rres, double* inertia, double* ep, double* ev,
                                           double* moment, unsigned int np);
/* reduction (decl:345): FP64 sum of np device doubles, written to *out (host). */
int xdem_reduction(double* in, unsigned int np, double* out);
/* cu_update_meshObjectData (decl:346-348) */
int xdem_cu_update_meshObjectData(double* vList, double* sph, double* dlocal, xdem_triangle_info* poly,
                                  double* dbi, unsigned int ntriangle);

/* ========================================================================== step layer */

typedef struct xdem_ctx xdem_ctx;

#define XDEM_PRECISION_FP64  0  /* everything in FP64 */
#define XDEM_PRECISION_MIXED 1  /* contact-force arithmetic in FP32; positions, Verlet state,
                                   tangential history and all force sums in FP64 */

typedef struct xdem_ctx_config {
    int precision;               /* XDEM_PRECISION_* */
    unsigned int p2p_capacity;   /* sphere-contact history slots per particle (0 = 12) */
    xdem_quirks quirks;
    unsigned int block_size;     /* threads per block of the contact kernel (0 = tuned default) */
    unsigned int grid_dim_max;   /* largest bucket count per axis (power of two, 0 = 512) */
    int fuse_velocity;           /* 1: rolling + velocity update inside the contact kernel;
                                    0: separate kernel (fewer registers in the contact kernel) */
    int device;                  /* HIP device ordinal */
    unsigned int buckets_per_sphere; /* bucket budget of a scene: the power of two at or above
                                        this times its sphere count (0 = 4); the largest axes
                                        are halved until the grid fits */
    double skin;                 /* neighbour-list skin [m]: the grid and the lists are rebuilt
                                    only on steps after a sphere has moved half of it since
                                    the last build (or after a state write or a wall move).
                                    0 = adaptive: between a sixteenth of skin_cap and
                                    skin_cap times the scene's smallest radius (2.5% and 40%
                                    by default), starting at a sixteenth, doubled when the lists
                                    lasted fewer than 16 steps and halved when they lasted
                                    more than 64; < 0 = no skin (rebuilt every step).
                                    Results do not depend on it. */
    int step_graph;              /* > 0: xdem_ctx_step replays a HIP graph of this many steps
                                    (made even) for runs at least that long, which saves the
                                    host's launch time and the gaps between kernels when steps
                                    are short (the full and lite contact passes stay concurrent
                                    branches). 0 (default): every kernel launched from the host.
                                    Results do not depend on it. */
    unsigned int rebuild_wait_limit; /* tries (~1 us each) a block of the list rebuild waits for
                                    a phase to complete before it gives up, sets the
                                    XDEM_ERR_GRID_SYNC warning and leaves the kernel (0 = 2^20,
                                    about a second); the blocks that stay complete the build,
                                    so results do not change. Small values exist to test that
                                    path. */
    unsigned int full_blocks;    /* blocks of the full-pass kernels (pair, triangle and contact
                                    passes), rounded down to a multiple of 64 (0 = 1024). They
                                    stride over the queued spheres, so results do not depend on
                                    it; a few blocks per CU keep several waves per SIMD to hide
                                    the neighbour gathers. */
    unsigned int merge_pairs_below; /* below this many spheres in a step's full pass the contact
                                    kernel walks their pairs itself instead of a pair pass of
                                    its own (0 = 65536; 1 = never). Results do not depend on
                                    it. */
    int keep_order;              /* 0 (default): the context stores each scene's spheres in the
                                    bucket order of their initial positions (a stable sort on
                                    the bucket key at commit), so that the neighbours a kernel
                                    gathers lie close in memory whatever order the caller gave
                                    them in; 1: in the caller's order. Every call that takes or
                                    returns per-sphere data (read, write, histories, part
                                    files) uses the caller's order either way, the reference's
                                    visiting order breaks its ties by the caller's ids, and
                                    results do not depend on it. Device pointers
                                    (xdem_ctx_device_ptr) address the stored order:
                                    xdem_ctx_sphere_order maps it. */
    double skin_cap;             /* the adaptive skin's largest value as a fraction of each
                                    scene's smallest radius (0 = 0.4); the skin adapts between
                                    a sixteenth of it and it, and buckets are sized for it. A
                                    larger cap means fewer list builds and more listed pairs
                                    (xdem_autotune tries 0.2, 0.4 and 0.8). Results do not
                                    depend on it. */
} xdem_ctx_config;

/* Fields for xdem_ctx_read / xdem_ctx_write / xdem_ctx_device_ptr (reference layouts). */
#define XDEM_FIELD_POS     0  /* double4 x,y,z,r */
#define XDEM_FIELD_VEL     1  /* double3 */
#define XDEM_FIELD_ACC     2  /* double3 */
#define XDEM_FIELD_EP      3  /* double4 */
#define XDEM_FIELD_EV      4  /* double4 */
#define XDEM_FIELD_EA      5  /* double4 */
#define XDEM_FIELD_FORCE   6  /* double3, valid after xdem_ctx_collide */
#define XDEM_FIELD_MOMENT  7  /* double3, valid after xdem_ctx_collide */
#define XDEM_FIELD_MASS    8  /* double */
#define XDEM_FIELD_INERTIA 9  /* double3 */
#define XDEM_FIELD_TMAX    10 /* double3 */
#define XDEM_FIELD_RRES    11 /* double */

/* History kinds for xdem_ctx_read_history. */
#define XDEM_HISTORY_P2P   0
#define XDEM_HISTORY_PLANE 1
#define XDEM_HISTORY_MESH  2  /* 5 slots; ids are triangle indices of the scene's mesh */
#define XDEM_HISTORY_CYLINDER 3 /* 3 slots; ids are cylinder ids (id + 1000: a lid seen from inside) */

/* A rigid triangle-mesh body (reference: xMeshObject + xParticleMeshObjectContact::define,
 * source/xdynamics_object/xParticleMeshObjectContact.cpp:77-167). Vertices are given in the
 * body frame; the pose places them (global = position + A(e) local). The proxy sphere of each
 * triangle (xUtilityFunctions::FitSphereToTriangle with factor 0.8, fitted to the global
 * triangle at this pose and kept in the body frame) fixes the reference's visiting order and
 * enlarges the scene's cell to twice the largest proxy radius
 * (xDiscreteElementMethodSimulation.cpp:91-100). A body with mass > 0 is dynamic: its
 * contact wrench is reduced every step for a multibody solver. */
typedef struct xdem_mesh {
    unsigned int ntriangle;
    const double* vertices;      /* [9*ntriangle] P, Q, R of every triangle, body frame */
    double dbi[15];              /* pose at definition, MM..ED3 */
    xdem_contact_property cp;    /* sphere-mesh material pair */
} xdem_mesh;

typedef struct xdem_ctx_info {
    unsigned int nscene, np, nsorted, nplane_total;
    unsigned int nbucket;        /* buckets of all scene grids */
    unsigned int key_bits;       /* bits of the largest bucket key */
    unsigned int p2p_capacity;
    unsigned long long device_bytes;
    unsigned long long steps_done;
    unsigned long long list_builds;  /* steps that rebuilt the grid and the neighbour lists */
    unsigned int list_spheres;       /* spheres with listed neighbours or a wall near (last build) */
    unsigned int contact_spheres;    /* spheres of the last step's full contact pass */
    unsigned int mesh_spheres;       /* spheres with triangles listed (last build) */
    unsigned int ntriangle;          /* triangles of all mesh bodies */
} xdem_ctx_info;

int xdem_ctx_create(const xdem_ctx_config* cfg, xdem_ctx** out);
int xdem_ctx_destroy(xdem_ctx* ctx);
/* Adds a scene; host arrays are copied. Scenes are independent (own walls, materials,
 * grid, time step). Must precede xdem_ctx_commit. */
int xdem_ctx_add_scene(xdem_ctx* ctx, const xdem_scene* scene, unsigned int* scene_index);
/* Allocates all device memory, uploads the scenes and sets the initial state of
 * xDiscreteElementMethodSimulation::Initialize (velocity as given, acc = (0, g.y, 0),
 * ep = (1,0,0,0), ev = ea = 0; xDiscreteElementMethodSimulation.cpp:67-80).
 * No step call allocates afterwards. */
int xdem_ctx_commit(xdem_ctx* ctx);
/* Advances every scene by nsteps full steps. */
int xdem_ctx_step(xdem_ctx* ctx, unsigned int nsteps);
/* One step in two halves: collide = position update, grid, all contacts (force, moment,
 * tmax, rres left in their arrays); finish = rolling resistance + velocity update. */
int xdem_ctx_collide(xdem_ctx* ctx);
int xdem_ctx_finish(xdem_ctx* ctx);
int xdem_ctx_synchronize(xdem_ctx* ctx);
/* Copies a field of one scene (its particles in input order) to/from host memory. */
int xdem_ctx_read(xdem_ctx* ctx, int field, unsigned int scene, double* host_out);
int xdem_ctx_write(xdem_ctx* ctx, int field, unsigned int scene, const double* host_in);
/* Device pointer of a field over all scenes (scenes concatenated in order), in the stored
 * sphere order (see keep_order and xdem_ctx_sphere_order). With fuse_velocity = 1,
 * XDEM_FIELD_VEL and XDEM_FIELD_EV alternate between two buffers from step to step: fetch the
 * pointer again after every step. */
int xdem_ctx_device_ptr(xdem_ctx* ctx, int field, void** out);
/* The stored order of a committed scene's spheres: order[j] (np values) is the caller's index
 * of the sphere stored j-th (row j of the scene's part of a device pointer). No counterpart in
 * the reference, which keeps its callers' order; it is for callers that use device pointers
 * with keep_order = 0. */
int xdem_ctx_sphere_order(xdem_ctx* ctx, unsigned int scene, unsigned int* order);
/* Re-sorts the stored order by the spheres' current positions (a stable sort on the bucket key,
 * each scene in its own range), for a pack that has mixed or settled since commit; the next
 * step rebuilds the lists. Call it between steps. Results do not depend on it. No counterpart
 * in the reference; it allocates and frees a temporary buffer (the widest per-sphere row
 * times the sphere count). */
int xdem_ctx_reorder(xdem_ctx* ctx);
/* History of one scene in the reference layout: count[np], ids[np*cap], tsd[2*np*cap]
 * (cap = p2p_capacity for XDEM_HISTORY_P2P, 3 for XDEM_HISTORY_PLANE, 5 for XDEM_HISTORY_MESH,
 * 3 for XDEM_HISTORY_CYLINDER).
 * Sphere ids are scene-local particle indices; sphere-pair slots past a count read as zeros. */
int xdem_ctx_read_history(xdem_ctx* ctx, int kind, unsigned int scene, unsigned int* count, unsigned int* ids,
                          double* tsd);
/* Writes one scene's history in the same layout (the inverse of xdem_ctx_read_history; the
 * restart of xContactManager::set_from_part_result, source/xdynamics_object/xContactManager.cpp
 * via xDynamicSimulator.cpp:468-487). A count above the slots, or a sphere or triangle id
 * outside the scene, refuses the whole history with XDEM_E_INVALID and writes nothing. */
int xdem_ctx_write_history(xdem_ctx* ctx, int kind, unsigned int scene, const unsigned int* count,
                           const unsigned int* ids, const double* tsd);
/* Device error word (XDEM_ERR_* bits) accumulated since the last clear. */
int xdem_ctx_error_word(xdem_ctx* ctx, unsigned int* out, int clear);
/* Body wrench of every wall of a scene from the last step: out[6*k .. 6*k+5] =
 * (Fx, Fy, Fz, Mx, My, Mz) exerted by the spheres on wall k (reference: six thrust
 * reductions per body, xParticlePlaneContact.cpp:159-168). */
int xdem_ctx_body_wrench(xdem_ctx* ctx, unsigned int scene, double* out);
/* New pose of wall k of a scene, 15 doubles in the MM..ED3 layout; the wall geometry is
 * moved with it (xParticlePlaneContact::update, xParticlePlaneContact.cpp:96-131). */
int xdem_ctx_set_body_pose(xdem_ctx* ctx, unsigned int scene, unsigned int wall, const double dbi[15]);
/* Adds the mesh body of a scene (at most one per scene; all of a scene's triangle contacts
 * share the reference's 5 history slots per sphere). Between add_scene and commit. */
int xdem_ctx_add_mesh(xdem_ctx* ctx, unsigned int scene, const xdem_mesh* mesh, unsigned int* mesh_index);
/* New pose of a mesh body (15 doubles, MM..ED3): its triangles and proxy centres are placed
 * from their body-frame coordinates (xParticleMeshObjectContact::update, impl:2544-2601). The
 * neighbour lists stay valid until the summed motion of the bodies since the last build
 * reaches half the skin. */
int xdem_ctx_set_mesh_pose(xdem_ctx* ctx, unsigned int mesh, const double dbi[15]);
/* Wrench (Fx, Fy, Fz, Mx, My, Mz) the spheres exerted on a dynamic mesh body in the last step
 * (-F and -(contact point - body position) x F summed over its contacts; the reference's six
 * reductions per body, xParticleMeshObjectContact.cpp:209-219). Zeros for a fixed mesh. */
int xdem_ctx_mesh_wrench(xdem_ctx* ctx, unsigned int mesh, double out[6]);
/* Adds a cylinder body to a scene (xParticleCylinderContact::define and its contact launch,
 * cu_cylinder_contact_force, decl:326-333 / impl:1327-1460): geometry as for
 * xdem_cu_cylinder_contact_force, pose dbi (15 doubles, MM..ED3), material pair cp. Every step
 * evaluates the scene's cylinders after its walls, in the order added, on XDEM_MAX_P2CY_COUNT
 * slots per sphere (a cylinder with id 0 restarts them, a later one appends). A cylinder with
 * mass > 0 is dynamic: its contact wrench is reduced every step. At most 16 per scene; between
 * add_scene and commit. The neighbour lists treat a cylinder whose axis ends are not on the body
 * frame's y axis, a length apart, as always in reach (every sphere of its scene takes the full
 * contact pass). */
int xdem_ctx_add_cylinder(xdem_ctx* ctx, unsigned int scene, const xdem_cylinder_info* cylinder,
                          const double dbi[15], const xdem_contact_property* cp, unsigned int* cylinder_index);
/* New pose of a cylinder body (15 doubles, MM..ED3; xParticleCylinderContact::update). A
 * rotation about its own axis leaves the neighbour lists valid. */
int xdem_ctx_set_cylinder_pose(xdem_ctx* ctx, unsigned int cylinder, const double dbi[15]);
/* Wrench (Fx, Fy, Fz, Mx, My, Mz) the spheres exerted on a dynamic cylinder in the last step;
 * zeros for a fixed one. */
int xdem_ctx_cylinder_wrench(xdem_ctx* ctx, unsigned int cylinder, double out[6]);
int xdem_ctx_info_get(xdem_ctx* ctx, xdem_ctx_info* out);
/* Sizes of one scene (valid after add_scene). */
typedef struct xdem_ctx_scene_info {
    unsigned int np, nplane, p2p_capacity, has_mesh;
    unsigned int first_particle;   /* its first particle in the context's concatenated arrays */
    unsigned int ncylinder;
} xdem_ctx_scene_info;
int xdem_ctx_scene_info_get(xdem_ctx* ctx, unsigned int scene, xdem_ctx_scene_info* out);
/* Part file of one scene: the step's result file of xResultManager::export_step_data_to_file
 * (source/xdynamics_manager/xResultManager.cpp:828-925, version 2 layout) -- ct, np, np,
 * mass[np], pos[4np], vel[3np], acc[3np], ep[4np], ev[4np], ea[4np], then the sphere, wall
 * (scenes with walls), cylinder (scenes with cylinders) and triangle (scenes with a mesh)
 * histories as count[np], ids[np cap], tsd[2 np cap]. load restores the state from such a file and returns its ct (the restart of
 * xDynamicsSimulator::set_from_part_result, xDynamicSimulator.cpp:392-492, reading the mass
 * block that the reference's reader skips); it keeps the contact histories as they are, as
 * the reference's restart does (xContactManager.cpp:316-322), unless flags holds
 * XDEM_PART_HISTORY, which restores them too. A file whose size does not fit the scene is
 * refused with XDEM_E_INVALID before anything is written. */
#define XDEM_PART_HISTORY 1
#define XDEM_PART_POSITIONS 2   /* only the positions (xParticleManager::SetCurrentParticlesFromPartResult,
                                   xParticleManager.cpp:563-644); not with XDEM_PART_HISTORY */
int xdem_ctx_save_part(xdem_ctx* ctx, unsigned int scene, const char* path, double ct);
int xdem_ctx_load_part(xdem_ctx* ctx, unsigned int scene, const char* path, int flags, double* ct);
/* Number of stored sphere-sphere and sphere-wall contacts of all scenes at the last step. */
int xdem_ctx_contact_count(xdem_ctx* ctx, unsigned long long* n_p2p, unsigned long long* n_wall);
/* hipStream_t the context works on. */
int xdem_ctx_stream(xdem_ctx* ctx, void** out);
/* Records nsteps full steps into a HIP graph (replaces any previous one). */
int xdem_ctx_graph_capture(xdem_ctx* ctx, unsigned int nsteps);
/* Launches the captured graph `nlaunch` times. */
int xdem_ctx_graph_launch(xdem_ctx* ctx, unsigned int nlaunch);
/* Phase times of the last xdem_ctx_step call, measured with HIP events on the context's
 * stream when timing is on (sums over the steps timed, at most 2048; steps are then
 * launched one by one, never from a graph):
 * out[0] = contact passes (after k_touch until both streams joined) ms, out[1] = k_rebuild
 * (the step's counters; bucket grid, sorted copies and neighbour lists on a rebuild) ms,
 * out[2] = number of steps timed, out[3] = rolling + velocity kernel ms (0 when fused
 * into the contact passes), out[4] = per-step pair test (k_touch) ms, out[5] = k_lite_move
 * ms (fused mode; 0 otherwise), out[6] = k_collide ms and out[7] = the triangle pass before it
 * (k_mesh_pass, with the wait for the fork; about 0 without meshes) ms, both on the second
 * stream, out[8] = the sphere-pair pass (k_pairs, on the second stream between the two; it
 * returns at once on a small full pass, whose pairs k_collide walks) ms. */
int xdem_ctx_set_timing(xdem_ctx* ctx, int on);
int xdem_ctx_timing(xdem_ctx* ctx, double out[9]);

/* ---------------------------------------------------------------- launch-shape tuner
 * No counterpart in the reference (its launch shapes are constants, decl:17 and
 * xDiscreteElementMethodSimulation.cpp:100). Times candidate launch shapes of the step layer
 * on the caller's scenes: for each candidate a context is built from `base` with that shape,
 * `warmup` steps run untimed and `steps` are timed; coordinate descent over the full-pass
 * blocks (256, 512, 1024, 2048) at 4 buckets per sphere, then the bucket budget (2, 4, 8),
 * then the adaptive skin's cap (0.2, 0.4, 0.8 of the smallest radius).
 * Results do not depend on the shapes, so only speed is compared; a candidate whose run sets
 * an error bit other than XDEM_ERR_GRID_SYNC is discarded. The winner is appended to
 * `cache_path` (a text file, one line per scene signature; NULL: no cache) and read back from
 * it on later calls with the same signature (no trials then, from_cache = 1). Use the result
 * as xdem_ctx_config.full_blocks / buckets_per_sphere / skin_cap (the cap is tried only when
 * base->skin is 0, the adaptive skin). Allocates and frees contexts: call it at set-up. */
#define XDEM_AUTOTUNE_MAX_TRIALS 16
typedef struct xdem_autotune_result {
    unsigned int full_blocks, buckets_per_sphere;
    double ms_per_step;             /* of the winner */
    int from_cache;
    unsigned int ntrial;
    unsigned int trial_full_blocks[XDEM_AUTOTUNE_MAX_TRIALS];
    unsigned int trial_buckets_per_sphere[XDEM_AUTOTUNE_MAX_TRIALS];
    double trial_ms_per_step[XDEM_AUTOTUNE_MAX_TRIALS];
    char key[256];                  /* the scene signature */
    double skin_cap;                /* of the winner (xdem_ctx_config.skin_cap) */
    double trial_skin_cap[XDEM_AUTOTUNE_MAX_TRIALS];
} xdem_autotune_result;
int xdem_autotune(const xdem_scene* scenes, unsigned int nscene, const xdem_ctx_config* base, unsigned int warmup,
                  unsigned int steps, const char* cache_path, xdem_autotune_result* out);

#ifdef __cplusplus
}
#endif
#endif

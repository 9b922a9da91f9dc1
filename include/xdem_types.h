/*
 * xdem_types.h -- plain-old-data types shared by the C ABI of the MI355X DEM step.
 *
 * Every struct here is byte-compatible with the structure of the same role in the
 * reference's launcher header, so a caller that fills the reference struct can hand
 * the same bytes to this library:
 *
 *   xdem_dem_parameters     <-> device_dem_parameters   include/xdynamics_parallel/xParallelDEM_decl.cuh:128-150
 *   xdem_contact_property   <-> device_contact_property include/xdynamics_parallel/xParallelDEM_decl.cuh:230-243
 *   xdem_plane_info         <-> device_plane_info       include/xdynamics_parallel/xParallelDEM_decl.cuh:161-174
 *   xdem_triangle_info      <-> device_triangle_info    include/xdynamics_parallel/xParallelDEM_decl.cuh:152-159
 *   xdem_cylinder_info      <-> device_cylinder_info    include/xdynamics_parallel/xParallelDEM_decl.cuh:220-228
 *   body pose double[15]    <-> MM..ED3 enum layout     include/xdynamics_parallel/xParallelDEM_decl.cuh:18-20
 *
 * Array arguments elsewhere in the ABI keep the reference's AoS layouts:
 * position double4 (x,y,z,radius), velocity/acceleration/force/moment double3,
 * Euler parameters and their rates double4 (e0,e1,e2,e3), mass double, inertia double3,
 * tangential history double2 (ds, s_dot_prev).
 */
#ifndef XDEM_TYPES_H
#define XDEM_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y, z; } xdem_double3;
typedef struct { unsigned int x, y, z; } xdem_uint3;

/* Global step parameters (reference: device_dem_parameters, decl:128-150). */
typedef struct xdem_dem_parameters {
    unsigned char rollingCondition; /* C++ bool in the reference: 1 byte + 3 padding */
    unsigned char _pad0[3];
    int contact_model;              /* 0 = DHS, 1 = Hertz-Mindlin no-slip (xTypes.h:166-167) */
    unsigned int np;
    unsigned int nmp;
    unsigned int nCluster;
    unsigned int nClusterObject;
    unsigned int nTsdaConnection;
    unsigned int nTsdaConnectionList;
    unsigned int nTsdaConnectionBodyData;
    unsigned int nsphere;
    unsigned int nplane;
    unsigned int ncylinder;
    unsigned int ncell;
    xdem_uint3 grid_size;
    double dt;
    double half2dt;
    double cell_size;
    double cohesion;
    xdem_double3 gravity;
    xdem_double3 world_origin;
} xdem_dem_parameters;

/* Per contact-object material data (reference: device_contact_property, decl:230-243). */
typedef struct xdem_contact_property {
    double Ei, Ej;
    double pri, prj;
    double Gi, Gj;
    double rest;
    double fric;
    double s_fric;
    double rfric;
    double coh;
    double sratio;
    double amp;
} xdem_contact_property;

/* Wall geometry (reference: device_plane_info, decl:161-174). */
typedef struct xdem_plane_info {
    unsigned int id;
    unsigned int _pad0;
    double l1, l2;
    xdem_double3 u1, u2, uw, xw, pa, pb, w2, w3, w4;
} xdem_plane_info;

/* Triangle of a mesh body (reference: device_triangle_info, decl:152-159). */
typedef struct xdem_triangle_info {
    int id;
    unsigned int tid;
    xdem_double3 P, Q, R;
} xdem_triangle_info;

/* Cylinder body (reference: device_cylinder_info, decl:220-228): a tube of zero wall thickness
 * about the body-frame segment pbase -> ptop, placed by the body pose. len_rr = (length, radius
 * of the side wall, radius of the inner test and the rim), which xParticleCylinderContact::define
 * fills from (cylinder_length, cylinder_bottom_radius, cylinder_top_radius). empty_part: 0 both
 * lids closed, 1 top lid open, 2 bottom lid open (xCylinderObject::empty_part). */
typedef struct xdem_cylinder_info {
    unsigned int id;
    unsigned int empty_part;
    double thickness;
    xdem_double3 len_rr;
    xdem_double3 pbase, ptop;
} xdem_cylinder_info;

/* Indices into the 15-double body pose (reference enum at decl:20). */
enum {
    XDEM_MM = 0, XDEM_PX, XDEM_PY, XDEM_PZ, XDEM_VX, XDEM_VY, XDEM_VZ,
    XDEM_E0, XDEM_E1, XDEM_E2, XDEM_E3, XDEM_ED0, XDEM_ED1, XDEM_ED2, XDEM_ED3
};

/* Switches for the reference's known behaviours that differ between its CPU and GPU
 * texts or look unintended. The default (all members as listed) is the parity profile:
 * CPU force law, correct Verlet, Ft carried across pairs, flat inertia indexing. */
typedef struct xdem_quirks {
    int gpu_force_law;        /* 0: damping of xContact.cpp:238-264; 1: damping of impl:432-452 */
    int velocity_not_updated; /* 0: Verlet of xIntegratorVV.cpp:143-149; 1: impl:207-221 keeps v */
    int stale_ft;             /* 1: Ft keeps the previous pair's value when |e| = 0 (impl:645) */
    int flat_inertia;         /* 1: rolling reads ((double*)inertia)[id] (impl:2437) */
} xdem_quirks;

/* History capacities of the reference (decl:13-15). */
#define XDEM_MAX_P2P_COUNT 12
#define XDEM_MAX_P2PL_COUNT 3
#define XDEM_MAX_P2MS_COUNT 5
#define XDEM_MAX_P2CY_COUNT 3

/* Device error word bits, readable after a step (the reference only printf's). */
#define XDEM_ERR_P2P_OVERFLOW   0x1u  /* more sphere contacts than history slots */
#define XDEM_ERR_PLANE_OVERFLOW 0x2u  /* more wall contacts than wall slots */
#define XDEM_ERR_MESH_OVERFLOW  0x4u  /* more triangle contacts than triangle slots */
#define XDEM_ERR_NONFINITE      0x8u  /* NaN/Inf in a force or a state update */
#define XDEM_ERR_MESH_LIST      0x10u /* more than 20 triangle candidates of one kind */
#define XDEM_ERR_NBR_OVERFLOW   0x20u /* more near neighbours than the pair-search list holds
                                         (step layer; the contacts beyond it are missed) */
#define XDEM_ERR_GRID_SYNC      0x40u /* warning (step layer): a block of the list rebuild stopped
                                         waiting for a phase (~1 s, or the configured limit); the
                                         blocks that stayed complete the build, results unchanged */
#define XDEM_ERR_BUILD_INCOMPLETE 0x80u /* (step layer) a list rebuild did not complete; the step ran
                                         without contacts. Not expected: the rebuild's phases always
                                         complete with the blocks that stay */
#define XDEM_ERR_WORKSPACE      0x100u /* (launcher layer) a launch needed more scratch than the
                                         workspace reserved for it; the launch did not run */
#define XDEM_ERR_MESH_NEAR      0x200u /* (step layer) more triangles within reach of a sphere than its
                                         triangle list holds (the contacts beyond it are missed) */
#define XDEM_ERR_CYL_OVERFLOW   0x400u /* more cylinder contacts than cylinder slots (the reference writes
                                         past them) */

/* Status codes returned by every C entry point. */
#define XDEM_OK              0
#define XDEM_E_INVALID      -1  /* bad argument (null pointer, size out of range) */
#define XDEM_E_HIP          -2  /* a HIP runtime call failed */
#define XDEM_E_NOT_READY    -3  /* call order violated (e.g. step before initialize) */
#define XDEM_E_DEVICE_ERROR -4  /* the device error word is non-zero */
#define XDEM_E_UNSUPPORTED  -5

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
static_assert(sizeof(xdem_dem_parameters) == 144, "device_dem_parameters layout");
static_assert(offsetof(xdem_dem_parameters, grid_size) == 52, "device_dem_parameters layout");
static_assert(offsetof(xdem_dem_parameters, dt) == 64, "device_dem_parameters layout");
static_assert(sizeof(xdem_contact_property) == 104, "device_contact_property layout");
static_assert(sizeof(xdem_plane_info) == 240, "device_plane_info layout");
static_assert(sizeof(xdem_triangle_info) == 80, "device_triangle_info layout");
static_assert(sizeof(xdem_cylinder_info) == 88, "device_cylinder_info layout");
#endif

#endif /* XDEM_TYPES_H */

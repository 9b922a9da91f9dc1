// xParallelDEM_decl.hpp -- drop-in replacement for the reference's launcher header
// include/xdynamics_parallel/xParallelDEM_decl.cuh (lines 128-174, 230-257, 280-348).
//
// Same type names, function names, argument order and argument meaning as the reference;
// every function forwards to the C ABI of libxdem.so (xdem_capi.h). The reference's
// launchers return void, so the shim does too and latches the first non-zero status in
// xdem_shim_status(), which the caller may inspect and clear (the reference checks errors
// with checkCudaErrors after the fact; the latch plays that role). reduction() returns the
// sum as the reference does; on failure it returns NaN and latches the status.
//
// Not declared here because they are outside the DEM step of spheres, planes and meshes:
// the cube, cylinder, cluster and spring-damper launchers (decl:310-315, 326-334, 350-424).
#ifndef XPARALLELDEM_DECL_HPP
#define XPARALLELDEM_DECL_HPP

#include <cmath>

#include "xdem_capi.h"

typedef xdem_dem_parameters device_dem_parameters;       // decl:128-150
typedef xdem_contact_property device_contact_property;   // decl:230-243
typedef xdem_plane_info device_plane_info;               // decl:161-174
typedef xdem_triangle_info device_triangle_info;         // decl:152-159
typedef xdem_cylinder_info device_cylinder_info;         // decl:220-228

// First non-zero status returned by a forwarded call since the last clear.
inline int& xdem_shim_status()
{
    static int status = 0;
    return status;
}

inline void xdem_shim_latch(int rc)
{
    if (rc != 0 && xdem_shim_status() == 0) xdem_shim_status() = rc;
}

// decl:280
inline void setDEMSymbolicParameter(device_dem_parameters* h_paras)
{
    xdem_shim_latch(xdem_setDEMSymbolicParameter(h_paras));
}

// decl:282-283
inline void vv_update_position(double* pos, double* ep, double* vel, double* ev, double* acc, double* ea,
                               unsigned int np)
{
    xdem_shim_latch(xdem_vv_update_position(pos, ep, vel, ev, acc, ea, np));
}

// decl:284-285
inline void vv_update_velocity(double* vel, double* acc, double* ep, double* ev, double* ea, double* force,
                               double* moment, double* mass, double* iner, unsigned int np)
{
    xdem_shim_latch(xdem_vv_update_velocity(vel, acc, ep, ev, ea, force, moment, mass, iner, np));
}

// decl:287
inline void cu_calculateHashAndIndex(unsigned int* hash, unsigned int* index, double* pos, unsigned int sid,
                                     unsigned int np)
{
    xdem_shim_latch(xdem_cu_calculateHashAndIndex(hash, index, pos, sid, np));
}

// decl:288-290
inline void cu_calculateHashAndIndexForPolygonSphere(unsigned int* hash, unsigned int* index, unsigned int sid,
                                                     unsigned int nsphere, double* sphere)
{
    xdem_shim_latch(xdem_cu_calculateHashAndIndexForPolygonSphere(hash, index, sid, nsphere, sphere));
}

// decl:291
inline void cu_reorderDataAndFindCellStart(unsigned int* hash, unsigned int* index, unsigned int* cstart,
                                           unsigned int* cend, unsigned int* sorted_index, unsigned int np,
                                           unsigned int ncell)
{
    xdem_shim_latch(xdem_cu_reorderDataAndFindCellStart(hash, index, cstart, cend, sorted_index, np, ncell));
}

// decl:293-299
inline void cu_calculate_p2p(double* pos, double* ep, double* vel, double* omega, double* force,
                             double* moment, double* mass, double* tmax, double* rres,
                             unsigned int* pair_count, unsigned int* pair_id, double* tsd,
                             unsigned int* sorted_index, unsigned int* cstart, unsigned int* cend,
                             device_contact_property* cp, unsigned int np)
{
    xdem_shim_latch(xdem_cu_calculate_p2p(pos, ep, vel, omega, force, moment, mass, tmax, rres, pair_count,
                                          pair_id, tsd, sorted_index, cstart, cend, cp, np));
}

// decl:302-308
inline void cu_plane_contact_force(device_plane_info* plan, double* dbi, device_contact_property* cp,
                                   double* pos, double* ep, double* vel, double* omega, double* force,
                                   double* moment, double* mass, double* tmax, double* rres,
                                   unsigned int* pair_count, unsigned int* pair_id, double* tsd,
                                   unsigned int np)
{
    xdem_shim_latch(xdem_cu_plane_contact_force(plan, dbi, cp, pos, ep, vel, omega, force, moment, mass, tmax,
                                                rres, pair_count, pair_id, tsd, np));
}

// decl:326-333
inline void cu_cylinder_contact_force(device_cylinder_info* cyl, double* bi, device_contact_property* cp, double* pos,
                                      double* ep, double* vel, double* omega, double* force, double* moment,
                                      double* mass, double* tmax, double* rres, unsigned int* pair_count,
                                      unsigned int* pair_id, double* tsd, unsigned int np)
{
    xdem_shim_latch(xdem_cu_cylinder_contact_force(cyl, bi, cp, pos, ep, vel, omega, force, moment, mass, tmax, rres,
                                                   pair_count, pair_id, tsd, np));
}

// decl:317-324
inline void cu_particle_polygonObject_collision(device_triangle_info* dpi, double* dbi, double* pos,
                                                double* ep, double* vel, double* omega, double* force,
                                                double* moment, double* mass, double* tmax, double* rres,
                                                unsigned int* pair_count, unsigned int* pair_id, double* tsd,
                                                double* dsph, unsigned int* sidx, unsigned int* cstart,
                                                unsigned int* cend, device_contact_property* cp,
                                                unsigned int np, unsigned int ntriangle)
{
    xdem_shim_latch(xdem_cu_particle_polygonObject_collision(dpi, dbi, pos, ep, vel, omega, force, moment, mass,
                                                             tmax, rres, pair_count, pair_id, tsd, dsph, sidx,
                                                             cstart, cend, cp, np, ntriangle));
}

// decl:336-343
inline void cu_decide_rolling_friction_moment(double* tmax, double* rres, double* inertia, double* ep,
                                              double* ev, double* moment, unsigned int np)
{
    xdem_shim_latch(xdem_cu_decide_rolling_friction_moment(tmax, rres, inertia, ep, ev, moment, np));
}

// decl:345
inline double reduction(double* in, unsigned int np)
{
    double out = 0.0;
    const int rc = xdem_reduction(in, np, &out);
    xdem_shim_latch(rc);
    return rc == 0 ? out : std::nan("");
}

// decl:346-348
inline void cu_update_meshObjectData(double* vList, double* sph, double* dlocal, device_triangle_info* poly,
                                     double* dbi, unsigned int np)
{
    xdem_shim_latch(xdem_cu_update_meshObjectData(vList, sph, dlocal, poly, dbi, np));
}

#endif  // XPARALLELDEM_DECL_HPP

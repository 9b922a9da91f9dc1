// xdem_cosim.cpp -- the DEM <-> multibody step loop of the reference's simulator for one mesh
// body driven by the host HHT integrator (xDynamicsSimulator::xRunSimulationThread,
// source/xdynamics_simulation/xDynamicSimulator.cpp:347-375): per step, the DEM step with the
// body at its current pose (the contact wrench of the step is reduced on the device), the HHT
// step of the body under that wrench (zeroed every step), and the new pose handed back to the
// DEM (xDiscreteElementMethodSimulation::updateObjectFromMBD).
#include "xdem_capi.h"
#include "xdem_mbd.h"

extern "C" int xdem_cosim_wheel_steps(xdem_ctx* ctx, xdem_hht_wheel* body, unsigned int mesh, unsigned int nsteps,
                                      double* track)
{
    if (!ctx || !body) return XDEM_E_INVALID;
    for (unsigned int k = 0; k < nsteps; k++) {
        int rc = xdem_ctx_step(ctx, 1);
        if (rc) return rc;
        double w[6], dbi[15];
        if ((rc = xdem_ctx_mesh_wrench(ctx, mesh, w))) return rc;
        int it = 0;
        if ((rc = xdem_hht_wheel_step(body, w, dbi, &it))) return rc;
        if ((rc = xdem_ctx_set_mesh_pose(ctx, mesh, dbi))) return rc;
        if (track) {   // per step: x, z, vx, vz, Fx, Fz, My, Newton iterations
            double* t = track + 8 * (size_t)k;
            t[0] = dbi[XDEM_PX]; t[1] = dbi[XDEM_PZ]; t[2] = dbi[XDEM_VX]; t[3] = dbi[XDEM_VZ];
            t[4] = w[0]; t[5] = w[2]; t[6] = w[4]; t[7] = (double)it;
        }
    }
    return XDEM_OK;
}

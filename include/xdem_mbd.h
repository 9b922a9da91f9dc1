/*
 * xdem_mbd.h -- host multibody side of the DEM <-> MBD coupling (plain C ABI, no GPU).
 *
 * The reference couples its DEM step to an implicit multibody integrator on the host: each
 * step zeroes the body forces, runs the DEM step (which adds the contact wrench of every
 * dynamic body), advances the bodies with HHT-alpha and hands their new poses back to the
 * DEM (xDynamicSimulator::xRunSimulationThread, source/xdynamics_simulation/xDynamicSimulator.cpp:347-375;
 * xIntegratorHHT::OneStepSimulation, source/xdynamics_simulation/xIntegratorHHT.cpp:101-183).
 *
 * xdem_hht_wheel is that integrator for one rigid body in Euler parameters, q = (x, y, z, e0..e3),
 * under the constraints of a driven wheel on a single-wheel test rig: y fixed, the body's y
 * axis kept along the global y axis, the rotation about y prescribed (theta = theta0 + omega t),
 * and e.e = 1. The wheel is free to move in x and z under gravity, a constant external load
 * and the DEM contact wrench. The HHT-alpha index-3 formulation, its prediction, Newton
 * correction and mass Jacobian follow xIntegratorHHT.cpp:101-202; alpha is applied after the
 * initial solve (the reference's Initialize resets it to -0.3, xIntegratorHHT.cpp:58).
 */
#ifndef XDEM_MBD_H
#define XDEM_MBD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xdem_hht_wheel xdem_hht_wheel;

typedef struct xdem_hht_wheel_config {
    double mass;              /* kg */
    double inertia[3];        /* principal moments in the body frame, kg m^2 */
    double dbi[15];           /* initial pose and rates, MM..ED3 layout (MM ignored) */
    double gravity[3];
    double load[3];           /* constant external force on the body, N */
    double omega;             /* prescribed rotation rate about global y, rad/s */
    double dt;
    double alpha;             /* HHT alpha in [-1/3, 0] */
    double tolerance;         /* Newton tolerance on the residual norm (0 = 1e-5, the reference's) */
    int max_iterations;       /* 0 = 100, the reference's limit */
} xdem_hht_wheel_config;

int xdem_hht_wheel_create(const xdem_hht_wheel_config* cfg, xdem_hht_wheel** out);
int xdem_hht_wheel_destroy(xdem_hht_wheel* w);
/* One HHT step from time t to t + dt with the contact wrench (Fx, Fy, Fz, Mx, My, Mz, global
 * frame, about the body position) of the DEM step just taken; the body's contact wrench is
 * zeroed every step (the GUI loop's semantics, xDynamicSimulator.cpp:363). Writes the new
 * pose in the MM..ED3 layout and the Newton iterations taken. */
int xdem_hht_wheel_step(xdem_hht_wheel* w, const double wrench[6], double dbi_out[15], int* iterations);
/* Current pose (MM..ED3), accelerations (7) and Lagrange multipliers (5: y, the two axis
 * constraints, the drive, the Euler-parameter norm). */
int xdem_hht_wheel_state(const xdem_hht_wheel* w, double dbi_out[15], double qdd_out[7], double lambda_out[5],
                         double* time);

/* The coupled loop for a mesh body of a step-layer context (xdem_capi.h), nsteps times:
 * xdem_ctx_step, the mesh's wrench, xdem_hht_wheel_step, xdem_ctx_set_mesh_pose. track (or
 * NULL) receives 8 doubles per step: x, z, vx, vz of the body, Fx, Fz, My of the wrench and
 * the Newton iterations. Implemented in libxdem.so. */
struct xdem_ctx;
int xdem_cosim_wheel_steps(struct xdem_ctx* ctx, xdem_hht_wheel* body, unsigned int mesh, unsigned int nsteps,
                           double* track);

#ifdef __cplusplus
}
#endif
#endif

/* xdem_calib.h -- gradients of a DEM run with respect to its contact material, for
 * calibration (no counterpart in the reference, which has no differentiation; the step it
 * differentiates is xIntegratorVV::OneStepSimulation, source/xdynamics_simulation/
 * xIntegratorVV.cpp:26-64, with the contact launchers of decl:293-308 and the rolling and
 * velocity kernels of impl:192-222 and impl:2420-2454).
 *
 * A calibration context holds one scene (spheres and walls, Hertz-Mindlin or DHS contact
 * without cohesion, the velocity-Verlet update of the reference). xdem_calib_forward runs it
 * from its initial state for nsteps FP64 steps at the parameters theta = (sliding friction,
 * restitution, s), where s scales both Young's moduli of every material by exp(s) (s = 0 is
 * the scene's material), storing the state every checkpoint_every steps. xdem_calib_backward
 * then takes the gradient of a scalar loss with respect to the final sphere positions and
 * returns d loss / d theta by the discrete adjoint of those steps: per segment of
 * checkpoint_every steps it restores the checkpoint, recomputes the steps while recording
 * their states and contacts, and sweeps them backwards (velocity update, rolling resistance,
 * every sphere-sphere and sphere-wall contact with its tangential history, position update),
 * each kernel's vector-Jacobian product formed from local Jacobians of the same code in
 * forward-mode dual numbers. Contact onset, the Coulomb cap and the slot bookkeeping are
 * taken as fixed by the recorded run (the derivative of the branch taken).
 *
 * The forward run is the reference's FP64 step operation for operation (a test compares it
 * bit for bit with the oracle). Limits: one scene, at most XDEM_CALIB_MAX_CAPACITY history
 * slots per sphere, XDEM_CALIB_TAPE contacts per sphere per step (more set
 * XDEM_ERR_P2P_OVERFLOW in the error word). Device memory: about (600 + 200) bytes per sphere
 * per step of a segment, plus a state per checkpoint. */
#ifndef XDEM_CALIB_H
#define XDEM_CALIB_H

#include "xdem_host.h"

#ifdef __cplusplus
extern "C" {
#endif

#define XDEM_CALIB_MAX_CAPACITY 16
#define XDEM_CALIB_TAPE 32

typedef struct xdem_calib xdem_calib;

typedef struct xdem_calib_config {
    unsigned int capacity;          /* sphere history slots per sphere (0 = 12, the reference's) */
    unsigned int checkpoint_every;  /* steps between stored states (0 = 100) */
    unsigned int max_steps;         /* longest run forward will be asked for (checkpoint storage) */
    int device;
    xdem_quirks quirks;             /* velocity_not_updated must be 0 */
} xdem_calib_config;

int xdem_calib_create(const xdem_scene* scene, const xdem_calib_config* cfg, xdem_calib** out);
int xdem_calib_destroy(xdem_calib* c);
/* Runs nsteps from the scene's initial state; pos_out (host, 4*np doubles, may be NULL)
 * receives the final positions and radii. */
int xdem_calib_forward(xdem_calib* c, const double theta[3], unsigned int nsteps, double* pos_out);
/* Continues from step `start` of the last run (a multiple of checkpoint_every, at most that
 * run's length) to nsteps at new parameters: the run is the last one's up to start and theta's
 * after it. xdem_calib_backward then sweeps only steps start..nsteps-1, giving the gradient
 * with respect to the parameters of that stretch: a finite-horizon gradient where the whole
 * run's is dominated by exploding sensitivities (contact onsets and slip reversals of a long
 * granular flow). */
int xdem_calib_forward_from(xdem_calib* c, const double theta[3], unsigned int start, unsigned int nsteps,
                            double* pos_out);
/* Reads the current state (host arrays of 3*np / 4*np doubles; NULL skips a field). */
int xdem_calib_state(xdem_calib* c, double* pos, double* vel, double* ep, double* ev);
/* dL_dpos: 3*np doubles (host), the loss gradient at the final positions of the last forward
 * run; grad receives d loss / d (friction, restitution, s) over that run's steps. */
int xdem_calib_backward(xdem_calib* c, const double* dL_dpos, double grad[3]);
/* Error word (XDEM_ERR_* bits) of the context's runs; clear != 0 resets it. */
int xdem_calib_error_word(xdem_calib* c, unsigned int* out, int clear);

/* The sphere-sphere contact of sphere i with j as one step evaluates it (host code, for
 * tests): inputs x_i, x_j, v_i, v_j, w_i, w_j (3 each), the history (ds, s_dot), theta;
 * out[12] = force, moment, rolling torque sum (3 each), rolling resistance, new (ds, s_dot);
 * jac[12*23] (may be NULL) their derivatives with respect to the 23 inputs in that order.
 * Returns 0 with out zeroed when the spheres do not overlap. */
int xdem_calib_pair_eval(const xdem_contact_property* cp, int contact_model, const xdem_quirks* q, double dt,
                         double ir, double jr, double im, double jm, const double in[23], double out[12],
                         double* jac);

#ifdef __cplusplus
}
#endif

#endif

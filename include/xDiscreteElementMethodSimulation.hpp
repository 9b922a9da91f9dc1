// xDiscreteElementMethodSimulation.hpp -- C++ facade of the MI355X DEM step with the
// method set of the reference's DEM entry points
// (include/xdynamics_simulation/xDiscreteElementMethodSimulation.h:13-37 and
// xIngegratorVV.h:13-14), built on the step layer of xdem_capi.h.
//
// Differences from the reference, by design:
//  - Initialize takes the scene(s) directly (the reference reads them from its object,
//    model and contact managers, which this build does not carry);
//  - several independent scenes may be advanced together;
//  - errors are returned as int status codes, and the device error word is checked by
//    OneStepSimulation's caller through ErrorWord().
#ifndef XDISCRETEELEMENTMETHODSIMULATION_HPP
#define XDISCRETEELEMENTMETHODSIMULATION_HPP

#include <vector>

#include "xdem_capi.h"

class xDiscreteElementMethodSimulation {
public:
    xDiscreteElementMethodSimulation();
    virtual ~xDiscreteElementMethodSimulation();

    // xDiscreteElementMethodSimulation::Initialize (xDiscreteElementMethodSimulation.cpp:35-246)
    virtual int Initialize(const xdem_scene* scenes, unsigned int nscene, const xdem_ctx_config* cfg = nullptr);
    bool Initialized() const;
    // Implemented by the integrator (xIntegratorVV::OneStepSimulation).
    virtual int OneStepSimulation(double ct, unsigned int cstep) = 0;
    // Copies the state of every scene to the host arrays (xDiscreteElementMethodSimulation.cpp:253-284).
    bool SaveStepResult(unsigned int pt);
    // xDiscreteElementMethodSimulation::CriticalTimeStep (…cpp:301-311) for scene 0's material.
    double CriticalTimeStep(double min_rad) const;
    // Pushes wall poses set with SetBodyPose to the device (…cpp:313-317).
    void updateObjectFromMBD();
    double* Position();   // device pointer, double4 per sphere, all scenes
    double* Velocity();   // device pointer, double3 per sphere
    unsigned int num_particles() const;
    unsigned int num_clusters() const;
    // Replaces the state of scene 0 from host arrays (…cpp:406-439); cpos is unused.
    unsigned int set_dem_data(double* cpos, double* pos, double* vel, double* acc, double* ep, double* ev, double* ea);

    // Boundary to a host multibody solver: per-wall wrench of the last step and new wall
    // poses (MM..ED3 layout) for the next step.
    int BodyWrench(unsigned int scene, double* out6_per_wall);
    int SetBodyPose(unsigned int scene, unsigned int wall, const double dbi[15]);
    unsigned int ErrorWord(bool clear = true);

    // Host copies filled by SaveStepResult (scene 0 order, reference layouts).
    const std::vector<double>& HostPosition() const { return h_pos; }
    const std::vector<double>& HostVelocity() const { return h_vel; }
    xdem_ctx* Context() { return ctx; }

protected:
    xdem_ctx* ctx;
    bool isInitilize;
    unsigned int np;
    std::vector<unsigned int> scene_np;
    std::vector<xdem_scene> scene_desc;
    std::vector<double> h_pos, h_vel, h_acc, h_ep, h_ev, h_ea;
    struct PendingPose { unsigned int scene, wall; double dbi[15]; };
    std::vector<PendingPose> pending;
};

// Velocity-Verlet integrator (xIntegratorVV.cpp:26-154).
class xIntegratorVV : public xDiscreteElementMethodSimulation {
public:
    xIntegratorVV();
    ~xIntegratorVV() override;
    int OneStepSimulation(double ct, unsigned int cstep) override;
};

#endif

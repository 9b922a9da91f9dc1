// xdem_simulation.cpp -- the C++ facade (xDiscreteElementMethodSimulation / xIntegratorVV)
// over the step layer of the C ABI.
#include "xDiscreteElementMethodSimulation.hpp"

#include <cstring>

xDiscreteElementMethodSimulation::xDiscreteElementMethodSimulation() : ctx(nullptr), isInitilize(false), np(0) {}

xDiscreteElementMethodSimulation::~xDiscreteElementMethodSimulation()
{
    if (ctx) xdem_ctx_destroy(ctx);
    ctx = nullptr;
}

int xDiscreteElementMethodSimulation::Initialize(const xdem_scene* scenes, unsigned int nscene, const xdem_ctx_config* cfg)
{
    if (!scenes || nscene == 0) return XDEM_E_INVALID;
    if (ctx) {
        xdem_ctx_destroy(ctx);
        ctx = nullptr;
    }
    isInitilize = false;
    // Position() and Velocity() hand out device pointers, which the reference's callers index
    // in their own particle order: the context keeps it (keep_order). Without a configuration,
    // xdem_ctx_create's defaults: FP64, the parity quirks.
    xdem_ctx_config local;
    if (cfg) {
        local = *cfg;
    } else {
        std::memset(&local, 0, sizeof(local));
        local.quirks.stale_ft = 1;
        local.quirks.flat_inertia = 1;
    }
    local.keep_order = 1;
    int rc = xdem_ctx_create(&local, &ctx);
    if (rc) return rc;
    np = 0;
    scene_np.clear();
    scene_desc.assign(scenes, scenes + nscene);
    for (unsigned s = 0; s < nscene; s++) {
        if ((rc = xdem_ctx_add_scene(ctx, &scenes[s], nullptr))) return rc;
        scene_np.push_back(scenes[s].np);
        np += scenes[s].np;
    }
    if ((rc = xdem_ctx_commit(ctx))) return rc;
    isInitilize = true;
    return XDEM_OK;
}

bool xDiscreteElementMethodSimulation::Initialized() const { return isInitilize; }

bool xDiscreteElementMethodSimulation::SaveStepResult(unsigned int pt)
{
    (void)pt;
    if (!isInitilize) return false;
    const size_t n = scene_np[0];
    h_pos.resize(4 * n); h_vel.resize(3 * n); h_acc.resize(3 * n);
    h_ep.resize(4 * n); h_ev.resize(4 * n); h_ea.resize(4 * n);
    return xdem_ctx_read(ctx, XDEM_FIELD_POS, 0, h_pos.data()) == XDEM_OK &&
           xdem_ctx_read(ctx, XDEM_FIELD_VEL, 0, h_vel.data()) == XDEM_OK &&
           xdem_ctx_read(ctx, XDEM_FIELD_ACC, 0, h_acc.data()) == XDEM_OK &&
           xdem_ctx_read(ctx, XDEM_FIELD_EP, 0, h_ep.data()) == XDEM_OK &&
           xdem_ctx_read(ctx, XDEM_FIELD_EV, 0, h_ev.data()) == XDEM_OK &&
           xdem_ctx_read(ctx, XDEM_FIELD_EA, 0, h_ea.data()) == XDEM_OK;
}

double xDiscreteElementMethodSimulation::CriticalTimeStep(double min_rad) const
{
    if (scene_desc.empty()) return 0.0;
    const xdem_scene& s = scene_desc[0];
    return xdem_critical_time_step(min_rad, s.density, s.cp_pp.Ei, s.cp_pp.pri);
}

void xDiscreteElementMethodSimulation::updateObjectFromMBD()
{
    for (const PendingPose& p : pending) xdem_ctx_set_body_pose(ctx, p.scene, p.wall, p.dbi);
    pending.clear();
}

double* xDiscreteElementMethodSimulation::Position()
{
    void* p = nullptr;
    if (ctx && xdem_ctx_device_ptr(ctx, XDEM_FIELD_POS, &p) == XDEM_OK) return static_cast<double*>(p);
    return nullptr;
}

double* xDiscreteElementMethodSimulation::Velocity()
{
    void* p = nullptr;
    if (ctx && xdem_ctx_device_ptr(ctx, XDEM_FIELD_VEL, &p) == XDEM_OK) return static_cast<double*>(p);
    return nullptr;
}

unsigned int xDiscreteElementMethodSimulation::num_particles() const { return np; }
unsigned int xDiscreteElementMethodSimulation::num_clusters() const { return np; }

unsigned int xDiscreteElementMethodSimulation::set_dem_data(double* cpos, double* pos, double* vel, double* acc, double* ep,
                                                            double* ev, double* ea)
{
    (void)cpos;
    if (!isInitilize) return 1;
    const struct { int f; const double* p; } w[6] = {{XDEM_FIELD_POS, pos}, {XDEM_FIELD_VEL, vel}, {XDEM_FIELD_ACC, acc},
                                                    {XDEM_FIELD_EP, ep},   {XDEM_FIELD_EV, ev},   {XDEM_FIELD_EA, ea}};
    for (const auto& x : w)
        if (x.p && xdem_ctx_write(ctx, x.f, 0, x.p) != XDEM_OK) return 1;
    return 0;
}

int xDiscreteElementMethodSimulation::BodyWrench(unsigned int scene, double* out)
{
    return ctx ? xdem_ctx_body_wrench(ctx, scene, out) : XDEM_E_NOT_READY;
}

int xDiscreteElementMethodSimulation::SetBodyPose(unsigned int scene, unsigned int wall, const double dbi[15])
{
    if (!dbi) return XDEM_E_INVALID;
    PendingPose p;
    p.scene = scene;
    p.wall = wall;
    std::memcpy(p.dbi, dbi, sizeof(p.dbi));
    pending.push_back(p);
    return XDEM_OK;
}

unsigned int xDiscreteElementMethodSimulation::ErrorWord(bool clear)
{
    unsigned int w = 0;
    if (ctx) xdem_ctx_error_word(ctx, &w, clear ? 1 : 0);
    return w;
}

xIntegratorVV::xIntegratorVV() {}
xIntegratorVV::~xIntegratorVV() {}

int xIntegratorVV::OneStepSimulation(double ct, unsigned int cstep)
{
    (void)ct;
    (void)cstep;
    if (!isInitilize) return XDEM_E_NOT_READY;
    return xdem_ctx_step(ctx, 1);
}
